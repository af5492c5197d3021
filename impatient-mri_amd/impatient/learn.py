"""Learned warm start and preconditioner gains for the Toeplitz CG.

Training data are (z0, x*) pairs the solver generates itself: for seeded
draws of a random ellipse phantom and a random field map on one trajectory,
the GPU synthesises k-space data (forward NUDFT + complex Gaussian noise),
forms b = F^H d, solves A x* = b with preconditioned CG to a tight
tolerance, and records the linear estimate z0 = M_chan^-1 b.

Learned parts:
  * the residual CNN (imp_cnn_*) x0 = z0 + s * net(z0 / s), trained on the
    GPU with its own forward / backward / Adam kernels;
  * radial gains g(|k|) of the circulant preconditioner, w = w_chan * g,
    the per-band least-squares fit of M^-1 A v ~ v over white probes v
    under the training field maps (g > 0, so M stays Hermitian positive
    definite and PCG stays valid); kept only when they beat the Chan weights
    on held-out draws.

At solve time the learned start is used only when its residual
||b - A x0|| is below ||b|| (the zero start's); otherwise CG starts from 0.

    python -m impatient.learn --config 3 --draws 200 --epochs 8 --out model.npz
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
import time
from dataclasses import dataclass

import numpy as np

from . import ops, synth
from ._lib import (CGStats, Context, IMP_CG_BEST_ITERATE, IMP_CG_FALLBACK_ZERO, IMP_CG_PRECOND,
                   IMP_CG_TRUE_RESIDUAL, IMP_CG_USE_X0, call, ptr)

NBINS = 64
# start from x0, or from 0 when ||b - A x0|| >= ||b||
START = IMP_CG_USE_X0 | IMP_CG_FALLBACK_ZERO


def random_phantom(rng, X, Y):
    """Random ellipse phantom: a body ellipse, 4-10 inner ellipses with
    signed intensities, values in about [0, 1.5]."""
    img = np.zeros(X.shape)

    def ell(A, a, b, x0, y0, ph):
        c, s = np.cos(ph), np.sin(ph)
        xr = (X - x0) * c + (Y - y0) * s
        yr = -(X - x0) * s + (Y - y0) * c
        return A * ((xr / a) ** 2 + (yr / b) ** 2 <= 1.0)

    a0, b0 = rng.uniform(0.30, 0.46, 2)
    img += ell(1.0, a0, b0, rng.uniform(-0.03, 0.03), rng.uniform(-0.03, 0.03),
               rng.uniform(0, np.pi))
    for _ in range(rng.integers(4, 11)):
        a, b = rng.uniform(0.02, 0.18, 2)
        x0, y0 = rng.uniform(-0.28, 0.28, 2)
        img += ell(rng.uniform(-0.5, 0.5), a, b, x0, y0, rng.uniform(0, np.pi))
    return np.clip(img, 0.0, None).astype(np.complex64)


class Geometry:
    """Device-resident trajectory, coils, Q_l and Toeplitz plan of one
    configuration; the field map changes per draw."""

    def __init__(self, ctx: Context, prob: synth.Problem):
        self.ctx, self.prob = ctx, prob
        up = lambda a: ctx.to_device(np.ascontiguousarray(a, dtype=np.float32))
        self.nk, self.nv, self.C = prob.nk, prob.nvox, prob.ncoils
        self.L = prob.ntime_segs
        self.kx, self.ky, self.t, self.x, self.y = (up(a) for a in (prob.kx, prob.ky, prob.t,
                                                                      prob.x, prob.y))
        self.fm = up(prob.fm)
        self.sr = up(prob.sens.real.reshape(-1)) if prob.sens is not None else None
        self.si = up(prob.sens.imag.reshape(-1)) if prob.sens is not None else None
        nq = 4 * self.nv
        self.qr = ctx.alloc(((self.L + 1) * nq,))
        self.qi = ctx.alloc(((self.L + 1) * nq,))
        call("imp_toeplitz_q", ctx.h, C.c_int(self.nk), ptr(self.kx), ptr(self.ky), None,
             ptr(self.t), None, None, C.c_int(prob.Nx), C.c_int(prob.Ny), C.c_int(1),
             C.c_int(nq), None, None, None, C.c_int(self.L), C.c_int(0), ptr(self.qr),
             ptr(self.qi))
        t32 = prob.t.astype(np.float32)
        self.tau = float(np.float32((t32[-1] - t32[0]) / np.float32(self.L))) if self.L else 0.0
        self.t0 = float(t32[0])
        h = C.c_void_p()
        call("imp_toep_create", ctx.h, C.c_int(prob.Ny), C.c_int(prob.Nx), C.c_int(1),
             C.c_int(self.C), C.c_int(self.L), C.c_float(self.tau), C.c_float(self.t0),
             ptr(self.qr), ptr(self.qi), ptr(self.fm), ptr(self.sr), ptr(self.si),
             C.c_float(prob.lam), C.byref(h))
        self.op = h
        self.chan = ctx.alloc((self.nv,))
        call("imp_toep_get_precond", h, ptr(self.chan))
        self.dr, self.di = ctx.alloc((self.C * self.nk,)), ctx.alloc((self.C * self.nk,))
        self.br, self.bi = ctx.alloc((self.nv,)), ctx.alloc((self.nv,))
        self.rr, self.ri = ctx.alloc((self.nv,)), ctx.alloc((self.nv,))

    def close(self):
        if self.op:
            call("imp_toep_destroy", self.op)
            self.op = None

    def set_fieldmap(self, fm):
        self.fm.upload(np.ascontiguousarray(fm, dtype=np.float32).reshape(-1))
        call("imp_toep_set_fieldmap", self.op, ptr(self.fm))

    def synth_data(self, rho, snr_db, rng=None, noise=None):
        """d = F rho (current field map) + noise at snr_db; leaves d on the device.
        The unit-variance noise is drawn from rng (real part, then imaginary
        part), or given as noise = (real, imaginary) drawn the same way."""
        rho = np.asarray(rho).reshape(-1)
        pr = self.ctx.to_device(np.ascontiguousarray(rho.real, np.float32))
        pi = self.ctx.to_device(np.ascontiguousarray(rho.imag, np.float32))
        call("imp_nudft_forward", self.ctx.h, C.c_int(self.nk), C.c_int(self.nv), C.c_int(self.C),
             ptr(self.kx), ptr(self.ky), None, ptr(self.t), ptr(self.x), ptr(self.y), None,
             ptr(self.fm), ptr(pr), ptr(pi), ptr(self.sr), ptr(self.si), ptr(self.dr),
             ptr(self.di))
        if snr_db is not None:
            dr, di = self.dr.get(), self.di.get()
            rms = np.sqrt(np.mean(dr.astype(np.float64) ** 2 + di.astype(np.float64) ** 2))
            sd = rms * 10 ** (-snr_db / 20) / np.sqrt(2)
            if noise is None:
                noise = (rng.standard_normal(dr.size, dtype=np.float32),
                         rng.standard_normal(di.size, dtype=np.float32))
            dr += (sd * noise[0]).astype(np.float32)
            di += (sd * noise[1]).astype(np.float32)
            self.dr.upload(dr)
            self.di.upload(di)

    def fhd(self):
        """b = F^H d into (br, bi)."""
        call("imp_nudft_adjoint", self.ctx.h, C.c_int(self.nk), C.c_int(self.nv), C.c_int(self.C),
             ptr(self.kx), ptr(self.ky), None, ptr(self.t), ptr(self.x), ptr(self.y), None,
             ptr(self.fm), None, None, ptr(self.dr), ptr(self.di), ptr(self.sr), ptr(self.si),
             ptr(self.br), ptr(self.bi))

    def precond(self, rr, ri, zr, zi):
        call("imp_toep_precond_apply", self.op, ptr(rr), ptr(ri), ptr(zr), ptr(zi))

    def solve(self, xr, xi, tol, max_iter=2000, flags=0):
        st = CGStats()
        t0 = time.perf_counter()
        call("imp_toep_cg", self.op, ptr(self.br), ptr(self.bi), ptr(xr), ptr(xi),
             C.c_int(max_iter), C.c_double(tol), C.c_int(flags), C.byref(st))
        d = st.as_dict()
        d["wall_s"] = time.perf_counter() - t0
        return d


@dataclass
class Pair:
    zr: object
    zi: object
    tr: object
    ti: object


def _draw_inputs(seed, X, Y, max_hz, n_noise):
    """Host side of one draw: phantom, field map and unit noise, in the order
    the seeded generator gives them."""
    rng = np.random.default_rng(seed)
    rho = random_phantom(rng, X, Y)
    fm = synth.field_map(rng, X, Y, 4, max_hz)
    noise = (rng.standard_normal(n_noise, dtype=np.float32),
             rng.standard_normal(n_noise, dtype=np.float32))
    return rho, fm, noise


def make_pairs(geo: Geometry, n: int, seed0: int, snr_db=40.0, tol=1e-6, log=print, ahead=3):
    """n training pairs (z0 = M_chan^-1 b, x*) on the device.  Worker threads
    draw the next `ahead` draws' phantoms, field maps and noise while the GPU
    solves the current one (the draws are the same as drawn inline)."""
    from concurrent.futures import ThreadPoolExecutor
    prob = geo.prob
    X = prob.x.reshape(prob.Ny, prob.Nx).astype(np.float64)
    Y = prob.y.reshape(prob.Ny, prob.Nx).astype(np.float64)
    pairs = []
    t0 = time.perf_counter()
    flags = IMP_CG_PRECOND | IMP_CG_BEST_ITERATE
    pool = ThreadPoolExecutor(max_workers=2)
    submit = lambda k: pool.submit(_draw_inputs, seed0 + k, X, Y, prob.meta["max_hz"],
                                   geo.C * geo.nk)
    pending = [submit(k) for k in range(min(ahead, n))]
    for k in range(n):
        rho, fm, noise = pending.pop(0).result()
        if k + ahead < n:
            pending.append(submit(k + ahead))
        geo.set_fieldmap(fm)
        geo.synth_data(rho, snr_db, noise=noise)
        geo.fhd()
        p = Pair(*(geo.ctx.alloc((geo.nv,)) for _ in range(4)))
        st = geo.solve(p.tr, p.ti, tol, flags=flags)
        geo.precond(geo.br, geo.bi, p.zr, p.zi)
        pairs.append(p)
        if (k + 1) % max(1, n // 10) == 0 or k == n - 1:
            log(f"pairs {k + 1}/{n}: last PCG {st['iterations']} it, "
                f"{time.perf_counter() - t0:.1f} s")
    pool.shutdown()
    return pairs


def train(net: ops.CNN, pairs, epochs, lr=2e-3, seed=0, log=print):
    """Adam over the pairs for `epochs` passes, cosine-decayed learning rate."""
    rng = np.random.default_rng(seed)
    total = epochs * len(pairs)
    step = 0
    hist = []
    t0 = time.perf_counter()
    for ep in range(epochs):
        order = rng.permutation(len(pairs))
        losses = []
        for i in order:
            p = pairs[i]
            lr_t = lr * 0.5 * (1 + np.cos(np.pi * step / total))
            losses.append(net.train_step_device(p.zr, p.zi, p.tr, p.ti, max(lr_t, lr * 1e-3)))
            step += 1
        hist.append(float(np.mean(losses)))
        log(f"epoch {ep + 1}/{epochs}: mean loss {hist[-1]:.4e} "
            f"({time.perf_counter() - t0:.1f} s)")
    return hist


def fit_radial_gains(geo: Geometry, pairs, nbins=NBINS):
    """Per radial band of |k/N|: g = sum Re(conj(Z0) X*) / sum |Z0|^2 over the
    training pairs (Z0 = DFT z0 = w_chan DFT b, X* = DFT x*), clipped to
    [0.05, 20] so every weight stays positive."""
    prob = geo.prob
    N1, N2 = prob.Ny, prob.Nx
    f1 = np.fft.fftfreq(N1)[:, None]
    f2 = np.fft.fftfreq(N2)[None, :]
    rho = np.sqrt(f1 ** 2 + f2 ** 2)
    rmax = 0.5 * np.sqrt(2.0)
    b = np.minimum(nbins - 1, (rho / rmax * nbins).astype(int)).reshape(-1)
    num = np.zeros(nbins)
    den = np.zeros(nbins)
    for p in pairs:
        Z = np.fft.fft2((p.zr.get() + 1j * p.zi.get()).reshape(N1, N2)).reshape(-1)
        Xs = np.fft.fft2((p.tr.get() + 1j * p.ti.get()).reshape(N1, N2)).reshape(-1)
        num += np.bincount(b, (np.conj(Z) * Xs).real, nbins)
        den += np.bincount(b, np.abs(Z) ** 2, nbins)
    g = np.where(den > 0, num / np.maximum(den, 1e-300), 1.0)
    return np.clip(g, 0.05, 20.0).astype(np.float32)


def probe_pairs(geo: Geometry, n: int, seed0: int, log=print):
    """Pairs (z = M_chan^-1 A v, v) for white complex Gaussian probes v under
    the field maps of n draws: M^-1 A ~ I per radial band is what a
    preconditioner needs, and unlike (M^-1 b, x*) these pairs carry no
    unconverged-solve or noise bias."""
    prob = geo.prob
    X = prob.x.reshape(prob.Ny, prob.Nx).astype(np.float64)
    Y = prob.y.reshape(prob.Ny, prob.Nx).astype(np.float64)
    pairs = []
    yr, yi = geo.ctx.alloc((geo.nv,)), geo.ctx.alloc((geo.nv,))
    for k in range(n):
        rng = np.random.default_rng(seed0 + k)
        geo.set_fieldmap(synth.field_map(rng, X, Y, 4, prob.meta["max_hz"]))
        p = Pair(*(geo.ctx.alloc((geo.nv,)) for _ in range(4)))
        p.tr.upload(rng.standard_normal(geo.nv).astype(np.float32))
        p.ti.upload(rng.standard_normal(geo.nv).astype(np.float32))
        call("imp_toep_apply", geo.op, ptr(p.tr), ptr(p.ti), ptr(yr), ptr(yi))
        geo.precond(yr, yi, p.zr, p.zi)
        pairs.append(p)
    log(f"{n} probe pairs for the preconditioner gains")
    return pairs


# The preconditioner network: a second instance of the same CNN, run on the
# frequency plane.  Input z (real, Im = 0): log w_chan(k), mean removed,
# fftshifted so DC sits at the image centre and the sampled disk is one
# connected region.  Output x = z + s net(z / s); the weights are
#   w(k) = w_chan(k) exp(clip(Re x(k) - z(k), -PC_CLIP, PC_CLIP)),
# real and positive, so M = F^H diag(1/w) F stays Hermitian positive
# definite and PCG stays valid whatever the network outputs.
PC_CLIP = 3.0


def precond_input(chan):
    """Network input from the Chan weights (natural DFT order, image shape):
    (z as complex64, the removed mean)."""
    lz = np.log(np.asarray(chan, np.float64))
    mu = float(lz.mean())
    return np.fft.fftshift(lz - mu).astype(np.complex64), mu


def precond_from_output(x, z, chan):
    """Weights from the network output x for input z (both fftshifted)."""
    d = np.clip(np.asarray(x).real - np.asarray(z).real, -PC_CLIP, PC_CLIP)
    return (np.asarray(chan, np.float64) * np.exp(np.fft.ifftshift(d))).astype(np.float32)


def precond_targets(geo: Geometry, n: int, probes: int, seed0: int, log=print):
    """Training targets of the preconditioner network, one per draw of a
    random field map: the per-frequency least-squares weight of
    M^-1 A v ~ v over `probes` white probes v,
      w(k) = sum Re(conj(AV(k)) V(k)) / sum |AV(k)|^2   (V = DFT v, AV = DFT A v),
    which is 1/lambda(k) for a circulant A; returned as the target image
    z + log(w / w_chan) (fftshifted), clipped to +-PC_CLIP.  Noisy per draw
    (a few probes per frequency): trained with a squared loss the network
    learns their mean, and its 3x3 stencils smooth it over frequency."""
    prob = geo.prob
    N1, N2 = prob.Ny, prob.Nx
    X = prob.x.reshape(N1, N2).astype(np.float64)
    Y = prob.y.reshape(N1, N2).astype(np.float64)
    chan = geo.chan.get().reshape(N1, N2).astype(np.float64)
    z, _ = precond_input(chan)
    yr, yi = geo.ctx.alloc((geo.nv,)), geo.ctx.alloc((geo.nv,))
    vr, vi = geo.ctx.alloc((geo.nv,)), geo.ctx.alloc((geo.nv,))
    out = []
    t0 = time.perf_counter()
    for k in range(n):
        rng = np.random.default_rng(seed0 + k)
        geo.set_fieldmap(synth.field_map(rng, X, Y, 4, prob.meta["max_hz"]))
        num = np.zeros((N1, N2))
        den = np.zeros((N1, N2))
        for _ in range(probes):
            v = (rng.standard_normal(geo.nv) + 1j * rng.standard_normal(geo.nv)).astype(np.complex64)
            vr.upload(v.real)
            vi.upload(v.imag)
            call("imp_toep_apply", geo.op, ptr(vr), ptr(vi), ptr(yr), ptr(yi))
            av = (yr.get() + 1j * yi.get()).reshape(N1, N2)
            V = np.fft.fft2(v.reshape(N1, N2))
            AV = np.fft.fft2(av)
            num += (np.conj(AV) * V).real
            den += np.abs(AV) ** 2
        w = np.maximum(num, 1e-30 * den.max()) / np.maximum(den, 1e-300)
        d = np.clip(np.log(w / chan), -PC_CLIP, PC_CLIP)
        out.append((z + np.fft.fftshift(d)).astype(np.complex64))
        if (k + 1) % max(1, n // 5) == 0 or k == n - 1:
            log(f"preconditioner targets {k + 1}/{n} ({probes} probes each), "
                f"{time.perf_counter() - t0:.1f} s")
    return z, out


def train_precond(net: ops.CNN, z, targets, epochs, lr=1e-3, seed=0, log=print):
    """Adam (the network's own kernels) on (z, target) for `epochs` passes."""
    ctx = net.ctx
    zr = ctx.to_device(np.ascontiguousarray(z.real.reshape(-1), np.float32))
    zi = ctx.to_device(np.ascontiguousarray(z.imag.reshape(-1), np.float32))
    tr, ti = ctx.alloc((z.size,)), ctx.alloc((z.size,))
    rng = np.random.default_rng(seed)
    total, step, hist = epochs * len(targets), 0, []
    for ep in range(epochs):
        losses = []
        for i in rng.permutation(len(targets)):
            tr.upload(np.ascontiguousarray(targets[i].real.reshape(-1), np.float32))
            ti.upload(np.ascontiguousarray(targets[i].imag.reshape(-1), np.float32))
            lr_t = lr * 0.5 * (1 + np.cos(np.pi * step / total))
            losses.append(net.train_step_device(zr, zi, tr, ti, max(lr_t, lr * 1e-3)))
            step += 1
        hist.append(float(np.mean(losses)))
    log(f"preconditioner network: {epochs} epochs over {len(targets)} targets, loss "
        f"{hist[0]:.3e} -> {hist[-1]:.3e}")
    return hist


def network_weights(net: ops.CNN, chan):
    """The preconditioner network's weights for an operator whose Chan
    weights are `chan` (natural DFT order, image shape)."""
    z, _ = precond_input(chan)
    return precond_from_output(net.forward(z), z, chan)


def learned_start(geo: Geometry, net: ops.CNN, xr, xi):
    """x0 = z0 + s * net(z0 / s) with z0 = M^-1 b (current preconditioner)."""
    geo.precond(geo.br, geo.bi, geo.rr, geo.ri)
    net.forward_device(geo.rr, geo.ri, xr, xi)


def evaluate(geo: Geometry, net: ops.CNN, gains, n, seed0, snr_db=40.0, tol=1e-5, log=print,
             pc_weights=None):
    """Held-out draws: iterations / wall time to tol for plain CG, Chan PCG,
    and the learned start with Chan or learned-gain preconditioning; with
    pc_weights (the preconditioner network's weights), PCG with those from
    zero ("netw") and from the learned start ("netw_learned")."""
    prob = geo.prob
    X = prob.x.reshape(prob.Ny, prob.Nx).astype(np.float64)
    Y = prob.y.reshape(prob.Ny, prob.Nx).astype(np.float64)
    xr, xi = geo.ctx.alloc((geo.nv,)), geo.ctx.alloc((geo.nv,))
    base = IMP_CG_BEST_ITERATE | IMP_CG_TRUE_RESIDUAL
    rows = []
    for k in range(n):
        rng = np.random.default_rng(seed0 + k)
        rho = random_phantom(rng, X, Y)
        fm = synth.field_map(rng, X, Y, 4, prob.meta["max_hz"])
        geo.set_fieldmap(fm)
        geo.synth_data(rho, snr_db, rng)
        geo.fhd()
        r = {}
        call("imp_toep_set_precond", geo.op, None)
        r["plain"] = geo.solve(xr, xi, tol, flags=base)
        r["chan"] = geo.solve(xr, xi, tol, flags=base | IMP_CG_PRECOND)
        # ablation: the linear estimate z0 = M^-1 b alone as the start
        geo.precond(geo.br, geo.bi, xr, xi)
        r["chan_linear"] = geo.solve(xr, xi, tol, flags=base | IMP_CG_PRECOND | START)
        t0 = time.perf_counter()
        learned_start(geo, net, xr, xi)
        geo.ctx.synchronize()
        t_start = time.perf_counter() - t0
        r["chan_learned"] = geo.solve(xr, xi, tol, flags=base | IMP_CG_PRECOND | START)
        r["chan_learned"]["start_s"] = t_start
        if gains is not None:
            gh = np.ascontiguousarray(gains, np.float32)
            call("imp_toep_set_precond_radial", geo.op, gh.ctypes.data_as(C.c_void_p),
                 C.c_int(len(gh)))
            r["gains"] = geo.solve(xr, xi, tol, flags=base | IMP_CG_PRECOND)
            # the network input is always the Chan estimate it was trained on
            call("imp_toep_set_precond", geo.op, None)
            t0 = time.perf_counter()
            learned_start(geo, net, xr, xi)
            call("imp_toep_set_precond_radial", geo.op, gh.ctypes.data_as(C.c_void_p),
                 C.c_int(len(gh)))
            geo.ctx.synchronize()
            t_start = time.perf_counter() - t0
            r["gains_learned"] = geo.solve(xr, xi, tol, flags=base | IMP_CG_PRECOND | START)
            r["gains_learned"]["start_s"] = t_start
            call("imp_toep_set_precond", geo.op, None)
        if pc_weights is not None:
            dw = geo.ctx.to_device(np.ascontiguousarray(np.asarray(pc_weights).reshape(-1),
                                                        np.float32))
            call("imp_toep_set_precond", geo.op, ptr(dw))
            r["netw"] = geo.solve(xr, xi, tol, flags=base | IMP_CG_PRECOND)
            call("imp_toep_set_precond", geo.op, None)
            learned_start(geo, net, xr, xi)
            call("imp_toep_set_precond", geo.op, ptr(dw))
            r["netw_learned"] = geo.solve(xr, xi, tol, flags=base | IMP_CG_PRECOND | START)
            call("imp_toep_set_precond", geo.op, None)
        rows.append(r)
        log("draw %d: " % k + ", ".join(f"{m} {v['iterations']} it" for m, v in r.items()))
    return rows


def save_model(path, params, gains, meta, params_pc=None):
    extra = {} if params_pc is None else {"params_pc": np.asarray(params_pc, np.float32)}
    np.savez(path, params=np.asarray(params, np.float32), gains=np.asarray(gains, np.float32),
             meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8), **extra)


def load_model(path):
    with np.load(path, allow_pickle=False) as z:
        meta = json.loads(bytes(z["meta"]).decode())
        return z["params"].astype(np.float32), z["gains"].astype(np.float32), meta


def load_precond_params(path):
    """The preconditioner network's parameters of a saved model, or None."""
    with np.load(path, allow_pickle=False) as z:
        return z["params_pc"].astype(np.float32) if "params_pc" in z.files else None


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--config", type=int, default=3)
    ap.add_argument("--draws", type=int, default=200)
    ap.add_argument("--val", type=int, default=6)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=None, help="base seed of the training draws")
    ap.add_argument("--pair-tol", type=float, default=1e-6,
                    help="relative residual the training solutions x* are solved to")
    ap.add_argument("--out", required=True)
    ap.add_argument("--precond-from", default=None, metavar="MODEL",
                    help="keep MODEL's warm-start network and gains; train only the "
                         "preconditioner network and write MODEL plus it to --out")
    ap.add_argument("--precond-draws", type=int, default=64)
    ap.add_argument("--precond-probes", type=int, default=4)
    ap.add_argument("--precond-epochs", type=int, default=40)
    args = ap.parse_args(argv)

    def log(m):
        print(f"[learn] {m}", flush=True)

    prob = synth.config(args.config)
    seed0 = args.seed if args.seed is not None else prob.meta["seed"] * 10
    ctx = Context(0)
    geo = Geometry(ctx, prob)
    if args.precond_from:
        return precond_main(args, ctx, geo, seed0, log)
    log(f"{prob.name}: {args.draws} training draws from seed {seed0}")
    pairs = make_pairs(geo, args.draws, seed0, tol=args.pair_tol, log=log)
    net = ops.CNN(ctx, prob.Ny, prob.Nx, seed=seed0)
    hist = train(net, pairs, args.epochs, args.lr, log=log)
    gains = fit_radial_gains(geo, probe_pairs(geo, 24, seed0 + 500000, log=log))
    log("radial gains: " + " ".join(f"{g:.3f}" for g in gains[::8]))
    # the preconditioner network
    net_pc = ops.CNN(ctx, prob.Ny, prob.Nx, seed=seed0 + 7)
    z, targets = precond_targets(geo, args.precond_draws, args.precond_probes, seed0 + 700000,
                                 log=log)
    hist_pc = train_precond(net_pc, z, targets, args.precond_epochs, log=log)
    w_net = network_weights(net_pc, geo.chan.get().reshape(prob.Ny, prob.Nx))
    rows = evaluate(geo, net, gains, args.val, seed0 + args.draws + 1000, log=log,
                    pc_weights=w_net)
    val = {m: [r[m]["iterations"] for r in rows] for m in rows[0]}
    # keep the learned gains only if they beat the Chan weights on held-out draws
    use_gains = float(np.mean(val["gains_learned"])) < float(np.mean(val["chan_learned"]))
    use_netw = float(np.mean(val["netw_learned"])) < min(float(np.mean(val["gains_learned"])),
                                                         float(np.mean(val["chan_learned"])))
    log(f"validation mean iterations: " + ", ".join(f"{m} {np.mean(v):.1f}" for m, v in val.items())
        + f"; learned gains {'kept' if use_gains else 'not used (Chan weights are better)'}; "
        f"network weights {'kept' if use_netw else 'not used'}")
    meta = dict(config=prob.name, draws=args.draws, epochs=args.epochs, lr=args.lr,
                pair_tol=args.pair_tol,
                seed0=seed0, loss_history=hist, nbins=NBINS, use_gains=use_gains,
                use_netw=use_netw, precond_draws=args.precond_draws,
                precond_probes=args.precond_probes, precond_epochs=args.precond_epochs,
                precond_loss_history=hist_pc, validation=val)
    save_model(args.out, net.get_params(), gains, meta, net_pc.get_params())
    log(f"saved {args.out}")
    net_pc.close()
    net.close()
    geo.close()
    ctx.close()
    return 0


def precond_main(args, ctx, geo, seed0, log):
    """--precond-from: add a preconditioner network to a trained model."""
    prob = geo.prob
    params, gains, meta = load_model(args.precond_from)
    net = ops.CNN(ctx, prob.Ny, prob.Nx, seed=1)
    net.set_params(params)
    net_pc = ops.CNN(ctx, prob.Ny, prob.Nx, seed=seed0 + 7)
    z, targets = precond_targets(geo, args.precond_draws, args.precond_probes, seed0 + 700000,
                                 log=log)
    hist_pc = train_precond(net_pc, z, targets, args.precond_epochs, log=log)
    w_net = network_weights(net_pc, geo.chan.get().reshape(prob.Ny, prob.Nx))
    rows = evaluate(geo, net, gains, args.val, seed0 + meta.get("draws", 0) + 1000, log=log,
                    pc_weights=w_net)
    val = {m: [r[m]["iterations"] for r in rows] for m in rows[0]}
    use_netw = float(np.mean(val["netw_learned"])) < min(float(np.mean(val["gains_learned"])),
                                                         float(np.mean(val["chan_learned"])))
    log("validation mean iterations: " + ", ".join(f"{m} {np.mean(v):.1f}" for m, v in val.items())
        + f"; network weights {'kept' if use_netw else 'not used'}")
    meta.update(use_netw=use_netw, precond_draws=args.precond_draws,
                precond_probes=args.precond_probes, precond_epochs=args.precond_epochs,
                precond_loss_history=hist_pc, precond_validation=val)
    save_model(args.out, params, gains, meta, net_pc.get_params())
    log(f"saved {args.out}")
    net_pc.close()
    net.close()
    geo.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
