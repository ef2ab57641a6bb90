"""The allocation-free apply (ncg_apply_noalloc) against ncg_apply, through the C ABI.

Two calls in a row on one device, the second with a smaller input (fewer load cases, and a
smaller problem), for an FMM operator (mixed precision: FP32 signatures, and FP64) and a
dense operator with more than 4 load cases (matrix-core path); each output equal to
ncg_apply's on the same inputs (the same kernels in the same order: bit for bit).  Device
buffers for x and y are allocated with the HIP runtime libncg itself links (through ctypes),
used here only as device memory.
"""
import ctypes

import numpy as np
import pytest

import ncref as R

pytestmark = pytest.mark.gpu


def _hip():
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    return hip


def _noalloc(ncg, prob, x):
    """y = A x through ncg_apply_noalloc with x and y in device memory"""
    hip = _hip()
    x = np.ascontiguousarray(x, np.complex128)
    y = np.zeros_like(x)
    xd, yd = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(xd), x.nbytes) == 0 and hip.hipMalloc(ctypes.byref(yd), y.nbytes) == 0
    try:
        assert hip.hipMemcpy(xd, x.ctypes.data, x.nbytes, 1) == 0          # host to device
        rc = ncg.load().ncg_apply_noalloc(prob._h, xd, yd, ctypes.c_int32(len(x)))
        if rc != 0:
            raise ncg.NcgError(rc, ncg.load().ncg_last_error().decode())
        assert hip.hipMemcpy(y.ctypes.data, yd, y.nbytes, 2) == 0          # device to host
    finally:
        hip.hipFree(xd)
        hip.hipFree(yd)
    return y


def _problem(ncg, name, method_bem):
    case = R.load_case(name)
    f = case["spec"]["freqs"][-1]
    pa = R.problem_arrays(case, f)
    prob = ncg.Problem(pa["nv"], pa["xyz"], pa["kind"], node_index=case["elems"], method_bem=method_bem)
    return prob, pa


def test_noalloc_apply_equals_apply():
    from mesh2hrtf_amd import ncg
    rng = np.random.default_rng(3)
    big, pb = _problem(ncg, "fmm_ml_head", 4)
    small, ps = _problem(ncg, "fmm_sl_point", 1)
    for prec in (ncg.MIXED, ncg.FP64):
        print("noalloc FMM precision", prec, flush=True)
        big.assemble(pb["k"], adm=pb["adm"], nrhs=3, precision=prec)
        small.assemble(ps["k"], adm=ps["adm"], nrhs=2, precision=prec)
        x1 = rng.standard_normal((3, pb["n"])) + 1j * rng.standard_normal((3, pb["n"]))
        x2 = rng.standard_normal((1, ps["n"])) + 1j * rng.standard_normal((1, ps["n"]))
        y1 = _noalloc(ncg, big, x1)
        y2 = _noalloc(ncg, small, x2)
        assert np.array_equal(y1, big.apply(x1))
        assert np.array_equal(y2, small.apply(x2))
    big.close()
    small.close()
    # dense operator, 20 then 6 load cases (matrix cores, FP64 accumulation)
    dense, pd = _problem(ncg, "soft_pw", 0)
    dense.assemble(pd["k"], adm=pd["adm"], nrhs=20, precision=ncg.MIXED)
    for nr in (20, 6):
        print("noalloc dense load cases", nr, flush=True)
        x = rng.standard_normal((nr, pd["n"])) + 1j * rng.standard_normal((nr, pd["n"]))
        assert np.array_equal(_noalloc(ncg, dense, x), dense.apply(x))
    dense.close()


def test_noalloc_rejects_oversized_and_bad_arguments():
    from mesh2hrtf_amd import ncg, project as P
    prob, pa = _problem(ncg, "soft_pw", 0)
    prob.assemble(pa["k"], adm=pa["adm"], nrhs=2)
    x = np.zeros((3, pa["n"]), np.complex128)
    with pytest.raises(ncg.NcgError):
        _noalloc(ncg, prob, x)                 # more load cases than assembled
    prob.close()
    # 64 load cases on the 20480-element head at 16 kHz in FP64: the signatures of its two
    # levels need about 436 MB, beyond the 64 MiB scratch
    mesh = P.config2_mesh(32)
    n = len(mesh.elems)
    big = ncg.Problem(np.full(n, 3), mesh.nodes[mesh.elems], np.zeros(n, np.int32), node_index=mesh.elems,
                      method_bem=4)
    src = P.ear_sources(mesh)[0]
    sp = P.point_source_spec([src] * 64, 16000.0)
    big.assemble(sp["k"], nrhs=64, point_sources=sp["point_sources"])
    with pytest.raises(ncg.NcgError, match="MiB of scratch"):
        _noalloc(ncg, big, np.zeros((64, n), np.complex128))
    big.close()
