"""imp_nudft_adjoint_noalloc: F^H d with every buffer in the library's fixed
per-GPU device array (no device allocation), called through the C ABI and
compared with imp_nudft_adjoint on the same inputs.

Two calls in a row with the second smaller (the array keeps the first
call's contents, so anything that must start from zero is cleared by the
call itself), the 1-2 coil (VALU) form, a 3D stack (summed directly), and
the size limit (rejected before any launch)."""
import ctypes as C

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL = 2e-6   # both entry points against each other (each is ~2e-7 from FP64)


def _adjoint(ctx, name, p, d, sens):
    from impatient._lib import call, ptr
    up = lambda a: ctx.to_device(np.ascontiguousarray(a, np.float32))
    d3 = p.Nz > 1
    kx, ky, t, x, y, fm = (up(a) for a in (p.kx, p.ky, p.t, p.x, p.y, p.fm))
    kz, z = (up(p.kz), up(p.z)) if d3 else (None, None)
    d = np.asarray(d).reshape(-1)
    dr, di = up(d.real), up(d.imag)
    sr = si = None
    if sens is not None:
        sr, si = up(np.asarray(sens).real.reshape(-1)), up(np.asarray(sens).imag.reshape(-1))
    nc = d.size // p.nk
    out_r, out_i = ctx.alloc((p.nvox,)), ctx.alloc((p.nvox,))
    opt = lambda a: ptr(a) if a is not None else None
    call(name, ctx.h, C.c_int(p.nk), C.c_int(p.nvox), C.c_int(nc), ptr(kx), ptr(ky), opt(kz),
         ptr(t), ptr(x), ptr(y), opt(z), ptr(fm), None, None, ptr(dr), ptr(di), opt(sr),
         opt(si), ptr(out_r), ptr(out_i))
    ctx.synchronize()
    return out_r.get() + 1j * out_i.get()


def _data(p, ncoils, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((ncoils, p.nk)) +
            1j * rng.standard_normal((ncoils, p.nk))).astype(np.complex64)


def _both(ctx, p, ncoils, seed):
    d = _data(p, ncoils, seed)
    sens = p.sens if ncoils > 1 else None
    a = _adjoint(ctx, "imp_nudft_adjoint_noalloc", p, d, sens)
    b = _adjoint(ctx, "imp_nudft_adjoint", p, d, sens)
    return rel_l2(a, b)


def test_noalloc_adjoint_matches_adjoint_twice_second_smaller(ctx):
    """64^2, 8 coils, 32 x 1024 samples (MFMA form, several chunks of
    partials in the array), then 32^2, 4 coils, 8 x 512."""
    from impatient import synth
    big = synth.tiny(N=64, ncoils=8, J=32, S=1024, ntime_segs=3, max_hz=150.0, seed=11)
    small = synth.tiny(N=32, ncoils=4, J=8, S=512, ntime_segs=3, max_hz=100.0, seed=12)
    e1 = _both(ctx, big, 8, 21)
    e2 = _both(ctx, small, 4, 22)
    print(f"noalloc vs adjoint: 64^2 8 coils {e1:.2e}; then 32^2 4 coils {e2:.2e}")
    assert e1 < TOL and e2 < TOL


def test_noalloc_adjoint_valu_form_and_3d(ctx):
    """One and two coils (per-lane VALU form), then a 3D stack of spirals."""
    from impatient import synth
    one = synth.tiny(N=32, ncoils=1, J=8, S=512, seed=13)
    two = synth.tiny(N=24, ncoils=2, J=8, S=256, seed=14)
    vol = synth.tiny(N=16, Nz=8, ncoils=4, J=8, S=256, seed=15)
    e = [_both(ctx, one, 1, 23), _both(ctx, two, 2, 24), _both(ctx, vol, 4, 25)]
    print("noalloc vs adjoint: 1 coil %.2e, 2 coils %.2e, 3D 4 coils %.2e" % tuple(e))
    assert max(e) < TOL


def test_noalloc_adjoint_rejects_what_does_not_fit(ctx):
    """512^2 with 32 coils and 48 x 8192 samples needs far more than the
    fixed array: the call fails with IMP_EUNSUPPORTED (before any launch,
    so it needs no data) and the next call still works."""
    from impatient import synth
    from impatient._lib import ImpError, call
    dummy = ctx.alloc((16,))
    p = dummy.ptr
    with pytest.raises(ImpError) as ei:
        call("imp_nudft_adjoint_noalloc", ctx.h, C.c_int(393216), C.c_int(262144), C.c_int(32),
             C.c_void_p(p), C.c_void_p(p), None, C.c_void_p(p), C.c_void_p(p), C.c_void_p(p),
             None, None, None, None, C.c_void_p(p), C.c_void_p(p), None, None, C.c_void_p(p),
             C.c_void_p(p))
    assert ei.value.status == -8   # IMP_EUNSUPPORTED
    assert "scratch" in str(ei.value)
    small = synth.tiny(N=16, ncoils=3, J=4, S=128, seed=16)
    assert _both(ctx, small, 3, 26) < TOL
