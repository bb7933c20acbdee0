"""The device's node-update operator (fmt2d_amd/csrc/eik_update.h) and the
host model build (fmt2d_amd/csrc/model.cpp), compiled for the CPU with g++
and iterated serially by tests/native/eik_cpu_check.cpp, against the
reference's own stage-1 fields (tests/golden/ref_fields_dx2.npz: a point
source in region 5, propagate2d.f90:33-127 after
initialize_region_from_source2d, :2307-2411).  This keeps a CPU-side check
on the exact operator the GPU kernels run: a change to eik_update.h that
moves a single bit of a traveltime or gradient fails here without a GPU.

The GPU tile schedule (emulated), plain Jacobi sweeps and in-place
Gauss-Seidel sweeps over the whole grid must all reach the
reference's field bit for bit (the fixed point does not depend on the
schedule)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from conftest import load_case

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "ref_fields_dx2.npz")


@pytest.fixture(scope="module")
def operator_lib(tmp_path_factory):
    d = tmp_path_factory.mktemp("eikcpu")
    so = str(d / "libeikcpu.so")
    srcs = [os.path.join(ROOT, "tests", "native", "eik_cpu_check.cpp"),
            os.path.join(ROOT, "fmt2d_amd", "csrc", "model.cpp")]
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", "-o", so] + srcs,
                   check=True)
    lib = C.CDLL(so)
    dp = C.POINTER(C.c_double)
    lib.cpu_model_create.restype = C.c_void_p
    lib.cpu_model_create.argtypes = [C.c_int, C.c_int, dp, dp, dp, dp, C.c_double, C.c_int, dp, dp, C.c_int, dp, dp,
                                     C.c_char_p, C.c_int]
    lib.cpu_region_nnode.argtypes = [C.c_void_p, C.c_int]
    lib.cpu_point_source.argtypes = [C.c_void_p, C.c_int, C.c_double, dp, dp, dp, C.c_int]
    return lib


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


@pytest.mark.parametrize("maxit", [0, -2000, 2000], ids=["tile_schedule", "jacobi", "gauss_seidel"])
def test_operator_reaches_reference_field(operator_lib, tmp_path, maxit):
    g = np.load(GOLD)
    for name in ("model.smesh", "moho.dat", "sed.dat"):
        with open(tmp_path / name, "w") as f:
            f.write(str(g["file_" + name.replace(".", "_")]))
    sm, sx, sz, mx, mz = load_case(str(tmp_path))
    vel = np.ascontiguousarray(sm.vel.reshape(-1))
    xs, tp, zr = (np.ascontiguousarray(a, dtype=np.float64) for a in (sm.xpos, sm.topo, sm.zrel))
    sx, sz, mx, mz = (np.ascontiguousarray(a, dtype=np.float64) for a in (sx, sz, mx, mz))
    err = C.create_string_buffer(256)
    h = operator_lib.cpu_model_create(len(xs), len(zr), _p(xs), _p(tp), _p(zr), _p(vel), sm.vwater, len(sx), _p(sx),
                                      _p(sz), len(mx), _p(mx), _p(mz), err, 256)
    assert h, err.value.decode()
    n = operator_lib.cpu_region_nnode(h, 5)
    for i, s in enumerate(g["sources"]):
        T, gx, gz = np.zeros(n), np.zeros(n), np.zeros(n)
        operator_lib.cpu_point_source(h, 5, float(s), _p(T), _p(gx), _p(gz), maxit)
        To = g[f"s{i}_tf1_T"]
        # unreached nodes: the reference's huge_time (1e20) and this build's
        # (1e20 as a float constant) differ in representation only
        fin = To < 1e19
        assert np.array_equal(fin, T < 1e19), f"source {s}: reached-node sets differ"
        assert np.array_equal(T[fin].view(np.int64), To[fin].view(np.int64)), \
            f"source {s}: {int(np.sum(T[fin] != To[fin]))} traveltimes differ"
        assert np.array_equal(gx[fin], g[f"s{i}_tf1_gx"][fin]), f"source {s}: dT/dx"
        assert np.array_equal(gz[fin], g[f"s{i}_tf1_gz"][fin]), f"source {s}: dT/dz"


def _model_handle(lib, d):
    sm, sx, sz, mx, mz = load_case(d)
    vel = np.ascontiguousarray(sm.vel.reshape(-1))
    xs, tp, zr = (np.ascontiguousarray(a, dtype=np.float64) for a in (sm.xpos, sm.topo, sm.zrel))
    sx, sz, mx, mz = (np.ascontiguousarray(a, dtype=np.float64) for a in (sx, sz, mx, mz))
    err = C.create_string_buffer(256)
    h = lib.cpu_model_create(len(xs), len(zr), _p(xs), _p(tp), _p(zr), _p(vel), sm.vwater, len(sx), _p(sx),
                             _p(sz), len(mx), _p(mx), _p(mz), err, 256)
    assert h, err.value.decode()
    return h


@pytest.mark.parametrize("maxit", [0, -2000, 2000], ids=["tile_schedule", "jacobi", "gauss_seidel"])
def test_fast_regular_path_matches_operator(operator_lib, tmp_path, plain_case, maxit):
    """reg_update (fmt2d_amd/csrc/eik_regular.h), the kernels' fast path for
    grid nodes with four regular cells, against update_grid_node on every
    state the iterations pass through (transient and converged, with and
    without second-order terms and equal-time ties): the same bits for T,
    both gradient components and the equal-time mask, at every evaluation."""
    lib = operator_lib
    lib.cpu_fast_check_set.argtypes = [C.c_int]
    lib.cpu_fast_check_get.argtypes = [C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    g = np.load(GOLD)
    for name in ("model.smesh", "moho.dat", "sed.dat"):
        with open(tmp_path / name, "w") as f:
            f.write(str(g["file_" + name.replace(".", "_")]))
    cases = [(str(tmp_path), [float(s) for s in g["sources"]]), (plain_case, [5.0, 95.0, 187.5])]
    total = 0
    for d, sources in cases:
        h = _model_handle(lib, d)
        n = lib.cpu_region_nnode(h, 5)
        for s in sources:
            T, gx, gz = np.zeros(n), np.zeros(n), np.zeros(n)
            lib.cpu_fast_check_set(1)
            lib.cpu_point_source(h, 5, s, _p(T), _p(gx), _p(gz), maxit if d != plain_case or maxit != -2000 else -400)
            checked, bad = C.c_longlong(), C.c_longlong()
            lib.cpu_fast_check_get(C.byref(checked), C.byref(bad))
            lib.cpu_fast_check_set(0)
            assert bad.value == 0, f"source {s}: {bad.value} of {checked.value} fast-path evaluations differ"
            total += checked.value
    assert total > 100000


def test_negative_discriminant_rule(operator_lib):
    """regular_update2d's negative discriminant (propagate2d.f90:745-756):
    taken as 0 while 0.5 sqrt(|det|)/|a| < 1e20, else the reference stops;
    the update then yields NaN, which fmt2d_forward reports as an error."""
    lib = operator_lib
    lib.cpu_reg_root.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double)]
    t = C.c_double()
    # det = b^2 - 4 a (c - 1) = 4 - 4*2*2 = -12 < 0: clamped, root -b / 2a
    assert lib.cpu_reg_root(2.0, -2.0, 3.0, 0.0, C.byref(t)) == 1 and t.value == 0.5
    # tmax above the root wins
    assert lib.cpu_reg_root(2.0, -2.0, 3.0, 0.75, C.byref(t)) == 1 and t.value == 0.75
    # 0.5 sqrt(4e-20 * 1e30) / 1e-20 = 1e25 >= 1e20: the reference's stop
    assert lib.cpu_reg_root(1e-20, 0.0, 1e30, 0.0, C.byref(t)) == 1 and np.isnan(t.value)
    # a == 0: no update
    assert lib.cpu_reg_root(0.0, 1.0, 1.0, 0.0, C.byref(t)) == 0
