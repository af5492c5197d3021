"""FP64 restatement of the oracle's Toeplitz operator and PCG in torch --
TEST INFRASTRUCTURE ONLY (a checker, like oracle/).

oracle.toeplitz_operator / oracle.pcg run the configuration-3 solve
(512^2, 32 coils, L = 10: 704 FFTs of 1024^2 points per apply) in minutes on
the host.  The same arithmetic in complex128 torch tensors runs on the GPU's
FP64 FFT (rocFFT through torch.fft) in seconds, so a converged FP64 solve can
sit in the default GPU suite.  It is the oracle's algorithm, not this
engine's: every function here follows the oracle function it names
(toeplitz/recon.cu:1113-1697 for the operator, textbook PCG for the solver)
and test_gpu_accuracy pins it against that oracle function on the same
inputs before using it.

complex64=True evaluates the operator's FFTs, Qf_l and products in
complex64 (FP32) instead, and pcg(..., store=np.complex64) rounds the solver
vectors to complex64 after every update: the two FP32 terms of the engine's
solve, separated.
"""
import numpy as np
import torch


def _fd_reg(p, lam, w=None):
    """oracle.fd_reg: lam * sum_d D_d^T W D_d p, periodic, on a [N1, N2(, N3)] tensor."""
    s = torch.zeros_like(p)
    for ax in range(p.ndim):
        if p.shape[ax] == 1:
            continue
        t1 = p - torch.roll(p, -1, dims=ax)
        if w is not None:
            t1 = t1 * w
        s = s + t1 - torch.roll(t1, 1, dims=ax)
    return lam * s


class ToeplitzOperator:
    """oracle.toeplitz_operator on the GPU: v -> sum_c conj(S_c) sum_l
    e^{+i w tau_l} T_l(e^{-i w tau_l} S_c v) + lam D^H W D v, with the
    low-corner embedding of the 2N grid and Qf_l = FFT(ifftshift(Q_l))."""

    def __init__(self, q, tau, t0, shape, fm=None, sens=None, lam=0.0, device="cuda",
                 complex64=False, coil_batch=8):
        self.cd = torch.complex64 if complex64 else torch.complex128
        rd = torch.float32 if complex64 else torch.float64
        N = tuple(shape) if len(shape) == 3 and shape[2] > 1 else tuple(shape[:2])
        self.N = N
        M = tuple(2 * n for n in N)
        self.M = M
        self.axes = tuple(range(1, len(N) + 1))
        qt = torch.as_tensor(np.asarray(q).reshape(q.shape[0], *M), dtype=torch.complex128,
                             device=device)
        # Qf_l in FP64 from the given Q_l, then stored in the operator's precision
        self.qf = torch.fft.fftn(torch.fft.ifftshift(qt, dim=self.axes), dim=self.axes).to(self.cd)
        f = torch.zeros(N, dtype=torch.float64, device=device) if fm is None else \
            torch.as_tensor(np.asarray(fm, np.float64).reshape(N), device=device)
        self.ph = [torch.exp(1j * f * (tau * l + t0)).to(self.cd) for l in range(qt.shape[0])]
        self.S = (torch.ones((1,) + N, dtype=self.cd, device=device) if sens is None else
                  torch.as_tensor(np.asarray(sens, np.complex128).reshape((-1,) + N),
                                  device=device).to(self.cd))
        self.lam = lam
        self.rd = rd
        self.device = device
        self.coil_batch = coil_batch

    def __call__(self, v):
        N, M = self.N, self.M
        pv = torch.as_tensor(v, device=self.device).reshape(N).to(self.cd)
        out = torch.zeros(N, dtype=self.cd, device=self.device)
        corner = (slice(None),) + tuple(slice(0, n) for n in N)
        for c0 in range(0, self.S.shape[0], self.coil_batch):
            Sc = self.S[c0:c0 + self.coil_batch]
            sp = Sc * pv
            acc = torch.zeros_like(sp)
            big = torch.zeros((Sc.shape[0],) + M, dtype=self.cd, device=self.device)
            for l in range(self.qf.shape[0]):
                big[corner] = torch.conj(self.ph[l]) * sp    # the rest of big stays zero
                X = torch.fft.fftn(big, dim=self.axes) * self.qf[l]
                r = torch.fft.ifftn(X, dim=self.axes)
                acc += self.ph[l] * r[corner]
            out += (torch.conj(Sc) * acc).sum(dim=0)
        if self.lam:
            out += _fd_reg(pv, self.lam)
        return out.reshape(-1).to(torch.complex128)


def circulant(w, device="cuda"):
    """oracle.circulant_apply: r -> IDFT(w DFT(r)) prod(dims), FP64."""
    wt = torch.as_tensor(np.asarray(w, np.float64), device=device)
    dims = tuple(range(wt.ndim))

    def apply(r):
        rr = torch.as_tensor(r, device=device).to(torch.complex128).reshape(wt.shape)
        return (torch.fft.ifftn(wt * torch.fft.fftn(rr, dim=dims), dim=dims) * wt.numel()).reshape(-1)
    return apply


def _vdot_re(a, b):
    return torch.sum(a.real * b.real + a.imag * b.imag).item()


def pcg(apply, precond, b, max_iter, tol=1e-16, best_iterate=True, store=np.complex128,
        device="cuda"):
    """oracle.pcg on tensors: rho = Re(r^H z), alpha = rho / Re(p^H A p),
    p = z + (rho'/rho) p; stop and best iterate on the unpreconditioned |r|.
    store=np.complex64 rounds x, r, z, p (and A p) to complex64 after each
    update; scalars stay FP64.  Returns (x as numpy complex128, iterations,
    residual history)."""
    if store == np.complex64:
        st = lambda v: v.to(torch.complex64).to(torch.complex128)
    else:
        st = lambda v: v
    b = st(torch.as_tensor(np.asarray(b, np.complex128).reshape(-1), device=device))
    x = torch.zeros_like(b)
    r = b.clone()
    z = st(precond(r))
    p = z.clone()
    rho = _vdot_re(r, z)
    normb = _vdot_re(b, b)
    best, xbest = normb, x.clone()
    hist = []
    it = 0
    while it < max_iter:
        q = st(apply(p))
        alpha = rho / _vdot_re(p, q)
        x = st(x + alpha * p)
        r = st(r - alpha * q)
        rr = _vdot_re(r, r)
        z = st(precond(r))
        rho_new = _vdot_re(r, z)
        p = st(z + (rho_new / rho) * p)
        rho = rho_new
        hist.append(float(np.sqrt(rr / normb)))
        it += 1
        if rr < best:
            best, xbest = rr, x.clone()
        if np.sqrt(rr / normb) < tol:
            break
    return (xbest if best_iterate else x).cpu().numpy(), it, hist
