"""numax.io's tensor files (.npy and NMX1) through the library
(nmx_tensor_file_*), checked against NumPy's own numpy.save / numpy.load --
an implementation of the format independent of this one -- and against the
reference's io tests (tests/io/test_npy.mojo, tests/io/test_io.mojo).

Header parsing runs without a GPU (nmx_tensor_file_info); every load or save
moves a payload to or from device memory and is marked gpu."""
import os
import struct

import numpy as np
import pytest

from numax_amd import _lib
from numax_amd.linalg import NumaxError

pytestmark_gpu = pytest.mark.gpu


def _info(path, fmt=0):
    import ctypes
    code, rank = ctypes.c_int(), ctypes.c_int()
    dims = (ctypes.c_int64 * 8)()
    st = _lib.load().nmx_tensor_file_info(os.fsencode(str(path)), fmt, ctypes.byref(code), ctypes.byref(rank), dims)
    return st, code.value, tuple(dims[i] for i in range(rank.value))


needs_lib = pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libnumax_hip.so not built")


@needs_lib
@pytest.mark.parametrize("arr,code", [(np.zeros((2, 3), np.float64), 2), (np.arange(5, dtype=np.float32), 1),
                                      (np.zeros((4, 1, 2), np.int32), 5), (np.zeros((3,), np.bool_), 11),
                                      (np.zeros((2, 2), np.uint8), 7), (np.float64(3.0), 2)])
def test_info_reads_numpy_written_headers(tmp_path, arr, code):
    p = tmp_path / "a.npy"
    np.save(p, arr)
    st, c, dims = _info(p)
    assert st == 0 and c == code and dims == np.shape(arr)


@needs_lib
def test_info_reads_a_version_2_header(tmp_path):
    p = tmp_path / "v2.npy"
    with open(p, "wb") as f:
        np.lib.format.write_array(f, np.ones((3, 4)), version=(2, 0))
    assert _info(p) == (0, 2, (3, 4))


@needs_lib
def test_info_errors(tmp_path):
    a = np.ones((2, 2))
    np.savez(tmp_path / "z.npz", a=a)
    assert _info(tmp_path / "z.npz")[0] == _lib.NMX_ERR_FORMAT  # a zip container
    np.save(tmp_path / "f.npy", np.asfortranarray(np.ones((2, 3))))
    assert _info(tmp_path / "f.npy")[0] == _lib.NMX_ERR_LAYOUT
    np.save(tmp_path / "b.npy", np.ones(3, dtype=">f8"))
    assert _info(tmp_path / "b.npy")[0] == _lib.NMX_ERR_LAYOUT
    np.save(tmp_path / "c.npy", np.ones(3, dtype=np.complex128))
    assert _info(tmp_path / "c.npy")[0] == _lib.NMX_ERR_DTYPE
    open(tmp_path / "short.npy", "wb").write(b"\x93NUM")
    assert _info(tmp_path / "short.npy")[0] == _lib.NMX_ERR_FORMAT
    assert _info(tmp_path / "missing.npy")[0] == _lib.NMX_ERR_IO


@needs_lib
def test_info_reads_an_nmx1_header(tmp_path):
    # NMX1 as numax/io/io.mojo:73-97 writes it, assembled byte by byte
    p = tmp_path / "a.nmx"
    name = b"float32"
    body = b"NMX1" + bytes([len(name)]) + name + bytes([2]) + struct.pack("<qq", 3, 5) + np.zeros(15, np.float32).tobytes()
    open(p, "wb").write(body)
    assert _info(p, 1) == (0, 1, (3, 5))
    open(p, "wb").write(b"NMX2" + body[4:])
    assert _info(p, 1)[0] == _lib.NMX_ERR_FORMAT


# ------------------------------------------------------------------ with a GPU
@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["float64", "float32", "int32", "uint8", "float16", "bool", "int64"])
@pytest.mark.parametrize("shape", [(7,), (3, 5), (2, 3, 4), (0, 4)])
def test_save_is_byte_identical_to_numpy_save(tmp_path, dtype, shape):
    """tests/io/test_npy.mojo:146-184."""
    import torch

    from numax_amd import io
    rng = np.random.default_rng(1)
    a = (rng.standard_normal(shape) * 100).astype(dtype)
    io.numpy.save(torch.as_tensor(a).cuda(), str(tmp_path / "ours.npy"))
    np.save(tmp_path / "theirs.npy", a)
    assert open(tmp_path / "ours.npy", "rb").read() == open(tmp_path / "theirs.npy", "rb").read()
    back = np.load(tmp_path / "ours.npy")
    assert back.dtype == a.dtype and back.shape == a.shape and np.array_equal(back, a)


@pytest.mark.gpu
@pytest.mark.parametrize("shape,dtype", [((5,), "float32"), ((3, 4), "int32"), ((40, 33), "float64"),
                                         ((2, 3, 4), "float64")])
def test_load_reads_numpy_written_files(tmp_path, shape, dtype):
    """tests/io/test_npy.mojo:122-144, 186-201."""
    import torch

    from numax_amd import io
    a = np.arange(int(np.prod(shape))).reshape(shape).astype(dtype) * 3 - 7
    np.save(tmp_path / "a.npy", a)
    got = io.numpy.load(str(tmp_path / "a.npy"), getattr(torch, dtype), shape)
    assert got.is_cuda and np.array_equal(got.cpu().numpy(), a)
    dyn = io.numpy.load_dyn(str(tmp_path / "a.npy"), getattr(torch, dtype))
    assert tuple(dyn.shape) == shape and np.array_equal(dyn.cpu().numpy(), a)
    with pytest.raises(NumaxError, match="rank mismatch"):
        io.numpy.load_dyn(str(tmp_path / "a.npy"), getattr(torch, dtype), rank=len(shape) + 1)


@pytest.mark.gpu
def test_load_errors_carry_numax_messages(tmp_path):
    """tests/io/test_npy.mojo:203-271."""
    import torch

    from numax_amd import io
    np.save(tmp_path / "a.npy", np.ones((2, 3)))
    with pytest.raises(NumaxError, match="dtype mismatch"):
        io.numpy.load(str(tmp_path / "a.npy"), torch.float32, (2, 3))
    with pytest.raises(NumaxError, match="shape mismatch"):
        io.numpy.load(str(tmp_path / "a.npy"), torch.float64, (3, 2))
    np.save(tmp_path / "f.npy", np.asfortranarray(np.ones((2, 3))))
    with pytest.raises(NumaxError, match="Fortran-ordered"):
        io.numpy.load(str(tmp_path / "f.npy"), torch.float64, (2, 3))
    open(tmp_path / "x.npy", "wb").write(b"not a numpy file at all")
    with pytest.raises(NumaxError, match="bad magic bytes"):
        io.numpy.load(str(tmp_path / "x.npy"), torch.float64, (2, 3))
    # a payload shorter than the header says
    raw = open(tmp_path / "a.npy", "rb").read()
    open(tmp_path / "t.npy", "wb").write(raw[:-8])
    with pytest.raises(NumaxError, match="shape mismatch"):
        io.numpy.load(str(tmp_path / "t.npy"), torch.float64, (2, 3))


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(4,), (3, 7), (2, 3, 5)])
def test_nmx_round_trip_and_errors(tmp_path, shape):
    """tests/io/test_io.mojo:39-136."""
    import torch

    from numax_amd import io
    a = torch.randn(shape, dtype=torch.float64, device="cuda")
    p = str(tmp_path / "a.nmx")
    io.nmx.save(a, p)
    raw = open(p, "rb").read()
    name = b"float64"
    assert raw[:4] == b"NMX1" and raw[4] == len(name) and raw[5:5 + len(name)] == name
    assert raw[5 + len(name)] == len(shape)
    assert struct.unpack(f"<{len(shape)}q", raw[6 + len(name):6 + len(name) + 8 * len(shape)]) == shape
    assert torch.equal(io.nmx.load(p, torch.float64, shape), a)
    with pytest.raises(NumaxError, match="dtype mismatch"):
        io.nmx.load(p, torch.float32, shape)
    with pytest.raises(NumaxError, match="shape mismatch"):
        io.nmx.load(p, torch.float64, shape[:-1] + (shape[-1] + 1,))
    with pytest.raises(NumaxError, match="shape mismatch"):  # rank
        io.nmx.load(p, torch.float64, shape + (1,))
    open(p, "wb").write(b"NMX0" + raw[4:])
    with pytest.raises(NumaxError, match="bad magic bytes"):
        io.nmx.load(p, torch.float64, shape)


@pytest.mark.gpu
def test_large_file_streams_through_both_staging_buffers(tmp_path):
    """A payload of several staging chunks (32 MiB each), ragged at the end."""
    import torch

    from numax_amd import io
    n = (5 * 32 * 2 ** 20) // 8 + 12345
    a = torch.arange(n, dtype=torch.float64, device="cuda") * 0.5 - 3.0
    p = str(tmp_path / "big.npy")
    io.numpy.save(a, p)
    back = np.load(p)
    assert back.shape == (n,) and back[0] == -3.0 and back[-1] == (n - 1) * 0.5 - 3.0
    assert np.array_equal(back[::99991], (np.arange(n, dtype=np.float64) * 0.5 - 3.0)[::99991])
    assert torch.equal(io.numpy.load(p, torch.float64, (n,)), a)


@pytest.mark.gpu
def test_npy_to_cholesky_flow(tmp_path):
    """examples/intermediate/npy_to_cholesky.mojo: NumPy writes an SPD matrix,
    numax loads it to the device and factors it, the factor goes back to a
    file NumPy reads, and it matches numpy.linalg.cholesky."""
    import torch

    from numax_amd import io
    from numax_amd import linalg as nl
    a = np.array([[4.0, 2.0, 0.6], [2.0, 5.0, 1.2], [0.6, 1.2, 3.0]])
    np.save(tmp_path / "spd.npy", a)
    l_dev = nl.cholesky(io.numpy.load(str(tmp_path / "spd.npy"), torch.float64, (3, 3)))
    io.numpy.save(l_dev, str(tmp_path / "l.npy"))
    np.testing.assert_allclose(np.load(tmp_path / "l.npy"), np.linalg.cholesky(a), rtol=0, atol=1e-15)
