"""Host-side checks (CPU only): the C-ABI library loads and exports every
declared symbol, the reference data formats round-trip, malformed inputs
return error codes, and the synthetic workloads have the stated layout."""
import ctypes as C
import re
import struct

import numpy as np
import pytest

from conftest import GOLDEN, REPO


def test_library_exports_every_declared_symbol():
    from impatient import _lib
    lib = _lib.load()
    header = (REPO / "include" / "impatient.h").read_text()
    names = sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(imp_\w+)\s*\(", header, re.M)))
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_status_and_version_strings():
    from impatient import _lib
    lib = _lib.load()
    assert b"gfx950" in lib.imp_version_string()
    assert lib.imp_status_string(-5) == b"malformed data file"
    assert lib.imp_status_string(0) == b"ok"


def _parse_v02(path):
    b = path.read_bytes()
    v, = struct.unpack("<f", b[:4])
    hdr = struct.unpack("<4i", b[4:20])
    return v, hdr, np.frombuffer(b[20:], dtype="<f4")


def test_read_reference_v02_files():
    from impatient import files
    d = GOLDEN / "ref_64x64x1"
    for name in ["kx", "ky", "t", "ix", "iy", "fm", "kdata_r", "kdata_i", "ixQ", "sensi_r"]:
        p = d / f"{name}.dat"
        data, info = files.read(p, with_info=True)
        v, hdr, raw = _parse_v02(p)
        assert abs(info.version - 0.2) < 1e-6
        assert (info.per_coil, info.ncoils, info.nslices, info.total) == hdr
        assert np.array_equal(data, raw)
    kx = files.read(d / "kx.dat")
    assert kx.size == 3770 and abs(kx[0] - 31.906708) < 1e-5
    ix = files.read(d / "ix.dat")
    # the shipped 2D image coordinates vary along y first (x constant over 64 entries)
    assert np.all(ix[:64] == ix[0]) and ix[64] != ix[0]


def _parse_v10(path):
    """Independent parser of the text-header format (structures.cpp:867-1110)."""
    b = path.read_bytes()
    pos, keys, data = 0, {}, []
    while True:
        end = b.index(b"\n", pos)
        line = b[pos:end].decode().strip()
        pos = end + 1
        if not line or line.startswith("//"):
            continue
        if line.startswith("EOF"):
            break
        if line == "Binary:":
            n = keys["Binary_Size"]
            data.append(np.frombuffer(b[pos:pos + 4 * n], dtype="<f4"))
            pos += 4 * n
            continue
        k, v = [s.strip() for s in line.split("=")]
        keys[k] = float(v) if k == "version" else int(v)
    return keys, np.concatenate(data)


@pytest.mark.parametrize("name,ncoils", [("t.dat", 1), ("sensi_i.dat", 4)])
def test_read_reference_v10_files(name, ncoils):
    from impatient import files
    p = GOLDEN / "ref_64x64x16_4coils" / name
    data, info = files.read(p, with_info=True)
    keys, raw = _parse_v10(p)
    assert info.version == 1.0 and info.ncoils == ncoils
    assert (info.nx, info.ny, info.nz) == (64, 64, 16)
    assert info.total == keys["file_size"] and info.per_coil == keys["file_size"] // ncoils
    assert np.array_equal(data, raw)


@pytest.mark.parametrize("ncoils", [-1, 1, 3])
def test_v10_round_trip(tmp_path, ncoils):
    from impatient import files
    rng = np.random.default_rng(0)
    a = rng.standard_normal(3 * 40).astype(np.float32)
    p = tmp_path / "a.dat"
    files.write_v10(p, a, 4, 5, 2, ncoils)
    b, info = files.read(p, with_info=True)
    assert np.array_equal(a, b)
    assert info.ncoils == (1 if ncoils == -1 else ncoils)
    keys, raw = _parse_v10(p)
    assert np.array_equal(raw, a)
    assert ("coil_number" in keys) == (ncoils != -1)


def test_v02_round_trip(tmp_path):
    from impatient import files
    a = np.arange(60, dtype=np.float32)
    p = tmp_path / "b.dat"
    files.write_v02(p, a, per_coil=20, ncoils=3)
    b, info = files.read(p, with_info=True)
    assert np.array_equal(a, b) and (info.per_coil, info.ncoils, info.total) == (20, 3, 60)
    v, hdr, raw = _parse_v02(p)
    assert abs(v - 0.2) < 1e-7 and hdr == (20, 3, 1, 60)


def _status(path):
    from impatient import _lib
    fi = _lib.FileInfo()
    buf = np.zeros(1024, np.float32)
    return _lib.load().imp_read_data_file(str(path).encode(), C.byref(fi),
                                          buf.ctypes.data_as(C.POINTER(C.c_float)), 1024)


def test_malformed_files_return_error_codes(tmp_path):
    assert _status(tmp_path / "missing.dat") == -3                 # IMP_EIO
    p = tmp_path / "short.dat"
    p.write_bytes(struct.pack("<f", 0.2) + struct.pack("<2i", 5, 1))
    assert _status(p) == -5                                          # truncated header
    p.write_bytes(struct.pack("<f", 0.2) + struct.pack("<4i", 5, 1, 1, 5) + b"\0" * 8)
    assert _status(p) == -3                                          # short data
    q = tmp_path / "v10.dat"
    q.write_bytes(b"version = 1.0\nfile_size = 2\nBinary_Size = 2\nBinary:\n" + b"\0" * 8)
    assert _status(q) == -5                                          # no EOF marker
    q.write_bytes(b"version = 1.0\nfile_size = 3\nBinary_Size = 2\nBinary:\n" + b"\0" * 8 + b"EOF\n")
    assert _status(q) == -5                                          # sections != file_size
    q.write_bytes(b"version = 1.0\nbogus = 2\nEOF\n")
    assert _status(q) == -5                                          # unknown key
    q.write_bytes(b"version = 1.0\nfile_size = 2\nBinary_Size = 2\nBinary:\n" + b"\0" * 8 + b"EOF\n")
    assert _status(q) == 0


def test_file_level_wrappers_reject_bad_arguments():
    """Argument errors are reported before any device work (no GPU needed)."""
    from impatient import _lib
    lib = _lib.load()
    argv = (C.c_char_p * 3)(b"toeplitz_recon", b"1", b"1")
    assert lib.imp_toeplitz_recon(3, argv, 0, 0, 1, 0, None, None, None, None, None) == -1
    argv18 = (C.c_char_p * 18)(*([b"1"] * 18))
    x = np.zeros(1, np.float32)
    P = x.ctypes.data_as(C.c_void_p)
    # -reg reads its inputs (here a missing directory) before any device work
    assert lib.imp_toeplitz_recon(18, argv18, 0, 1, 0, 0, None, P, P, None, None) == -3
    assert lib.imp_toeplitz_computeQ_GPU(b"/nonexistent", C.c_float(2), 8, 8, 1, None,
                                         None, None, 1, 0, C.c_float(1.0), 0) == -1
    # gridding reads its input files first: a missing directory is IMP_EIO
    assert lib.imp_toeplitz_computeFH_GPU(b"/nonexistent", C.c_float(2), 8, 8, 1, None,
                                          P, P, 0, 1, C.c_float(1.0)) == -3
    assert lib.imp_toeplitz_computeFH_GPU(b"/nonexistent", C.c_float(2), 8, 8, 1, None,
                                          P, P, 0, 0, C.c_float(1.0)) == -1


def test_context_without_gpu_fails_cleanly():
    import os
    from impatient import _lib
    if os.path.exists("/dev/kfd") and os.environ.get("IMP_ASSUME_NO_GPU") is None:
        pytest.skip("a GPU may be present")
    h = C.c_void_p()
    st = _lib.load().imp_ctx_create(0, C.byref(h))
    assert st == -6 and not h.value
    assert "no HIP device" in _lib.last_error()


def test_synthetic_workload_layout():
    from impatient import synth
    p = synth.config(0)
    assert (p.Nx, p.Ny, p.Nz, p.ncoils, p.nk) == (64, 64, 1, 1, 8192)
    assert p.kx.min() >= -32 and p.kx.max() < 32
    # x fastest: consecutive voxels step in x
    assert p.x[1] - p.x[0] == pytest.approx(1 / 64) and p.y[1] == p.y[0]
    assert np.abs(p.fm).max() == pytest.approx(2 * np.pi * 100, rel=1e-5)
    p1 = synth.make_problem("s", 16, 4, 4, 4, 64, 5e-3, 100.0, 3, 2, 1)
    ss = np.sum(np.abs(p1.sens) ** 2, axis=0)
    assert np.allclose(ss, 1.0, atol=1e-5)
    assert p1.z[1] - p1.z[0] == pytest.approx(1 / 4)        # z fastest in 3D
    assert sorted(set(p1.kz.tolist())) == [-2, -1, 0, 1]


def _write_mtx(path, rows, cols, I, J, V, banner="%%MatrixMarket matrix coordinate real general"):
    with open(path, "w") as f:
        f.write(banner + "\n% a comment line\n%\n")
        f.write(f"{rows} {cols} {len(V)}\n")
        for i, j, v in zip(I, J, V):
            f.write(f"{i + 1} {j + 1} {v:.17g}\n")


def test_read_mtx(tmp_path):
    """Matrix Market coordinate files through the C ABI (host only): sizes,
    1-based to 0-based indices, values rounded to float32 as readMtxFile
    stores them (kernels/mmio.cpp:760-770); scipy reads the same file."""
    import scipy.io
    from impatient import ops
    rng = np.random.default_rng(3)
    rows, cols, nnz = 12, 7, 30
    I = rng.integers(0, rows, nnz)
    J = rng.integers(0, cols, nnz)
    V = rng.standard_normal(nnz)
    f = tmp_path / "c.rmtx"
    _write_mtx(f, rows, cols, I, J, V)
    r, c, gi, gj, gv = ops.read_mtx(f)
    assert (r, c) == (rows, cols)
    assert np.array_equal(gi, I) and np.array_equal(gj, J)
    assert np.array_equal(gv, V.astype(np.float32))
    m = scipy.io.mmread(str(f)).tocoo()
    assert m.shape == (rows, cols) and np.allclose(m.toarray().sum(), V.sum())


def test_read_mtx_rejects_unsupported_and_bad_files(tmp_path):
    from impatient import _lib, ops
    cases = {
        "complex.rmtx": ("%%MatrixMarket matrix coordinate complex general", "3 3 1\n1 1 1 0\n"),
        "pattern.rmtx": ("%%MatrixMarket matrix coordinate pattern general", "3 3 1\n1 1\n"),
        "array.rmtx": ("%%MatrixMarket matrix array real general", "3 3\n"),
        "range.rmtx": ("%%MatrixMarket matrix coordinate real general", "3 3 1\n4 1 1.0\n"),
        "short.rmtx": ("%%MatrixMarket matrix coordinate real general", "3 3 2\n1 1 1.0\n"),
        "nobanner.rmtx": ("3 3 1", "1 1 1.0\n"),
    }
    for name, (banner, body) in cases.items():
        f = tmp_path / name
        f.write_text(banner + "\n" + body)
        with pytest.raises(_lib.ImpError) as e:
            ops.read_mtx(f)
        assert e.value.status in (-5, -1), name
    with pytest.raises(_lib.ImpError) as e:
        ops.read_mtx(tmp_path / "missing.rmtx")
    assert e.value.status == -3


def test_preconditioner_network_weights_stay_positive_and_bounded():
    """The preconditioner network's input and output maps (host side): the
    input is log w_chan, mean removed, with DC moved to the image centre;
    an output equal to its input gives back Chan's weights exactly; any
    output gives positive weights within exp(+-PC_CLIP) of Chan's, so the
    preconditioner M = F^H diag(1/w) F stays Hermitian positive definite."""
    from impatient import learn
    rng = np.random.default_rng(4)
    chan = np.exp(rng.uniform(-4, 2, size=(32, 48)))
    z, mu = learn.precond_input(chan)
    assert abs(np.log(chan).mean() - mu) < 1e-12
    # DC (index 0, 0 in natural order) sits at the centre after the shift
    assert abs(z[16, 24].real - (np.log(chan[0, 0]) - mu)) < 1e-5
    assert np.allclose(learn.precond_from_output(z, z, chan), chan, rtol=1e-6)
    big = z + 50.0 * rng.standard_normal(z.shape).astype(np.float32)
    w = learn.precond_from_output(big, z, chan).astype(np.float64)
    r = np.log(w / chan)
    assert np.all(w > 0) and np.all(np.isfinite(w))
    assert r.max() <= learn.PC_CLIP + 1e-5 and r.min() >= -learn.PC_CLIP - 1e-5
    assert r.max() > learn.PC_CLIP - 1e-3 and r.min() < -learn.PC_CLIP + 1e-3
    # a change at one shifted position moves the weight of the matching
    # natural-order frequency only
    x = z.copy()
    x[16 + 3, 24 - 5] += 1.0
    w1 = learn.precond_from_output(x, z, chan).astype(np.float64)
    moved = np.argwhere(np.abs(w1 / chan - 1) > 1e-6)
    assert moved.tolist() == [[3, -5 % 48]]
