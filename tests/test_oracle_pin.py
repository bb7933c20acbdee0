"""Pins of the C oracle (oracle/fmm2d_oracle.c).

* against the reference's own modules: golden stage fields of a coarse model
  (tests/golden/ref_fields_dx2.npz, written by tests/golden/make_ref_fields.py
  from oracle/_ref/ref_fields) must match bit for bit, and, where the
  reference build is present, the plain 401x61 case is re-run live;
* against closed forms: straight-ray times in a homogeneous half-space, the
  analytic first-arrival time of a constant-gradient medium, and the
  eikonal residual of the returned gradients, so that a dropped term or a wrong index fails here too.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_case

GOLD = os.path.join(ROOT, "tests", "golden", "ref_fields_dx2.npz")


def _golden_case(tmp_path):
    g = np.load(GOLD)
    for name in ("model.smesh", "moho.dat", "sed.dat"):
        with open(tmp_path / name, "w") as f:
            f.write(str(g["file_" + name.replace(".", "_")]))
    return g, load_case(str(tmp_path))


def test_golden_geometry(tmp_path):
    from oracle.pyoracle import OracleModel
    g, (sm, sx, sz, mx, mz) = _golden_case(tmp_path)
    om = OracleModel(sm, sx, sz, mx, mz)
    assert np.array_equal(om.fully_regular(), g["fully_regular"])
    assert np.array_equal(om.node_region(), g["node_region"])
    for r in range(1, 7):
        nodes = om.region_nodes(r)
        for key in ("i1", "i2", "i3", "x", "z", "vel"):
            assert np.array_equal(nodes[key], g[f"r{r}_{key}"]), (r, key)
    for k in range(1, 6):
        iv = om.isec(k)
        for key in iv:
            assert np.array_equal(iv[key], g[f"i{k}_{key}"]), (k, key)


def test_golden_stage_fields_bit_exact(tmp_path):
    from oracle.pyoracle import OracleModel
    g, (sm, sx, sz, mx, mz) = _golden_case(tmp_path)
    om = OracleModel(sm, sx, sz, mx, mz)
    for i, s in enumerate(g["sources"]):
        out = om.forward_source(float(s), 12, headwave=False)
        for itf in range(1, 13):
            T, gx, gz = out[itf - 1]
            assert np.array_equal(T, g[f"s{i}_tf{itf}_T"], equal_nan=True), (s, itf)
            assert np.array_equal(gx, g[f"s{i}_tf{itf}_gx"], equal_nan=True), (s, itf)
            assert np.array_equal(gz, g[f"s{i}_tf{itf}_gz"], equal_nan=True), (s, itf)


@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "oracle", "_ref", "ref_fields")),
                    reason="reference build (oracle/_ref) not present")
def test_live_reference_plain_case(plain_case, tmp_path):
    from oracle.pyoracle import OracleModel
    from oracle.refdump import read_refdump
    srcs = [5.0, 95.0]
    with open(os.path.join(plain_case, "ref_fields.in"), "w") as f:
        f.write("model.smesh\nmoho.dat\nsed.dat\nref.bin\n2\n5.0\n95.0\n12\n")
    subprocess.run([os.path.join(ROOT, "oracle", "_ref", "ref_fields")], cwd=plain_case, check=True,
                   capture_output=True)
    ref = read_refdump(os.path.join(plain_case, "ref.bin"), 2, 12)
    om = OracleModel(*load_case(plain_case))
    for i, s in enumerate(srcs):
        out = om.forward_source(s, 12, headwave=False)
        for itf, T, gx, gz in ref["sources"][i]:
            o = out[itf - 1]
            assert np.array_equal(o[0], T, equal_nan=True), (s, itf)
            assert np.array_equal(o[1], gx, equal_nan=True) and np.array_equal(o[2], gz, equal_nan=True)


# ---------------------------------------------------------------- closed forms

def _flat_model(v_of_z, z_sf=3.0, dx=0.5, xmax=100.0):
    """Flat seafloor on a grid row; velocity v_of_z(depth below sea level)."""
    from fmt2d_amd import formats
    nx = int(round(xmax / dx)) + 1
    x = np.arange(nx) * dx
    topo = np.full(nx, z_sf)
    zrel = np.arange(0.0, 22.0 + 1e-9, dx)
    vel = np.array([[v_of_z(z_sf + zr) for zr in zrel] for _ in range(nx)])
    sm = formats.Smesh(x, topo, zrel, vel, 1.5, 0.33)
    sed = z_sf + 1.0 + 0 * x
    moho = z_sf + 10.0 + 0 * x
    return sm, x, sed, x, moho


def test_homogeneous_axis_and_reciprocity():
    from oracle.pyoracle import OracleModel
    v = 5.0
    sm, sx, sz, mx, mz = _flat_model(lambda z: v)
    om = OracleModel(sm, sx, sz, mx, mz)
    nodes = om.region_nodes(5)
    T = om.forward_source(50.0, 1)[0][0]
    col = (nodes["i1"] == 101) & (nodes["z"] > 3.6)       # below the source
    assert np.allclose(T[col], (nodes["z"][col] - 3.0) / v, atol=1e-12)
    # off-node source: the mixed-order stencil stays within a few percent
    # of r/v away from the source
    T2 = om.forward_source(50.25, 1)[0][0]
    g = nodes["i1"] > 0
    r = np.hypot(nodes["x"][g] - 50.25, nodes["z"][g] - 3.0)
    rel = (T2[g] - r / v) / (r / v)
    far = r > 10.0
    assert np.median(np.abs(rel[far])) < 0.005 and np.abs(rel[far]).max() < 0.02
    # eikonal residual of the returned gradient: |grad T| v = 1
    full = om.forward_source(50.25, 1)[0]
    gn = np.hypot(full[1], full[2])[g][far] * v
    assert np.median(np.abs(gn - 1.0)) < 1e-3 and np.percentile(np.abs(gn - 1.0), 99) < 0.05


def test_constant_gradient_first_arrival():
    """v = v0 + g (z - zs): T = arccosh(1 + g^2 r^2 / (2 v0 v)) / g."""
    from oracle.pyoracle import OracleModel
    v0, grad, zs, src = 3.0, 0.25, 3.0, 50.25
    sm, sx, sz, mx, mz = _flat_model(lambda z: v0 + grad * (z - zs))
    om = OracleModel(sm, sx, sz, mx, mz)
    nodes = om.region_nodes(5)
    T = om.forward_source(src, 1)[0][0]
    g = (nodes["i1"] > 0) & (nodes["z"] > zs + 1.0)
    x, z = nodes["x"][g], nodes["z"][g]
    r = np.hypot(x - src, z - zs)
    vz = v0 + grad * (z - zs)
    exact = np.arccosh(1.0 + grad ** 2 * r ** 2 / (2.0 * v0 * vz)) / grad
    rel = (T[g] - exact) / exact
    far = r > 10.0
    assert abs(np.median(rel[far])) < 6e-3 and np.abs(rel[far]).max() < 3e-2


def test_headwave_post_pass_bit_exact(tmp_path):
    """The PmP head-wave post-pass of run_forward2d (inversion2d.f90:4294-4409)
    as the oracle restates it (pmp_headwave in oracle/fmm2d_oracle.c): tf5 of
    two OBS of the 0.5 km model against the reference inversion's own saved
    field (tests/golden/ref_headwave.npz), every traveltime and gradient bit
    for bit; without the post-pass the same field differs, so the pin
    reaches it."""
    from oracle.pyoracle import OracleModel
    g = np.load(os.path.join(ROOT, "tests", "golden", "ref_headwave.npz"))
    for name in ("model.smesh", "moho.dat", "sed.dat"):
        with open(tmp_path / name, "w") as f:
            f.write(str(g["file_" + name.replace(".", "_")]))
    om = OracleModel(*load_case(str(tmp_path)))
    changed = 0
    for k, s in enumerate(g["obs"]):
        with_hw = om.forward_source(float(s), 12, headwave=True)
        without = om.forward_source(float(s), 12, headwave=False)
        for itf in (1, 5):
            T, gx, gz = with_hw[itf - 1]
            assert np.array_equal(T, g[f"s{k}_tf{itf}_T"], equal_nan=True), (s, itf)
            assert np.array_equal(gx, g[f"s{k}_tf{itf}_gx"], equal_nan=True), (s, itf)
            assert np.array_equal(gz, g[f"s{k}_tf{itf}_gz"], equal_nan=True), (s, itf)
        T0, gx0, gz0 = without[4]
        changed += int(np.sum((T0 != g[f"s{k}_tf5_T"]) | (gx0 != g[f"s{k}_tf5_gx"]) | (gz0 != g[f"s{k}_tf5_gz"])))
    assert changed > 0
