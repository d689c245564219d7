"""The ABI from C: tests/c_driver.c (built by psphray_amd/build.sh with gcc
against include/psphray_amd.h) runs context -> tables (rates file +
kernel.akline) -> physics -> tree -> state -> rays -> one whole tick ->
download, on a 16^3 field; the result is checked against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import psphray_amd as psr
from psphray_amd.fields import make_workload
from oracle import pyoracle as po

from conftest import GOLDEN, RATES, ROOT, print_case

DRIVER = os.path.join(ROOT, "tests", "_bin", "c_driver")
F32 = ("sml", "mass", "rho", "ie", "yeMET")
F64 = ("lambdaH", "lambdaHeI", "lambdaHeII")
ACC = ("yGdepHI", "yGdepHeI", "yGdepHeII", "yGrecHII", "yGrecHeII", "yGrecHeIII", "heat")


def write_case(path, wl, rays, tick):
    f = wl.physics.flags
    ps = wl.ps
    with open(path, "wb") as o:
        np.array([ps.npar, len(rays), tick], np.int64).tofile(o)
        np.array([wl.boxsize, wl.tick_time, wl.physics.units.C_HMF], np.float64).tofile(o)
        np.array([f.isothermal, f.adiabatic, f.on_the_spot, f.h_only, f.secondary_ionization,
                  f.trace_only, f.no_photon, f.comoving], np.int32).tofile(o)
        np.array([f.redshift, f.fake_temperature], np.float64).tofile(o)
        np.ascontiguousarray(ps.pos, np.float32).tofile(o)
        for n in F32:
            np.ascontiguousarray(getattr(ps, n), np.float32).tofile(o)
        np.ascontiguousarray(ps.flag, np.int8).tofile(o)
        for n in F64:
            np.ascontiguousarray(getattr(ps, n), np.float64).tofile(o)
        for n in ACC:
            np.ascontiguousarray(getattr(ps, n), np.float32).tofile(o)
        np.ascontiguousarray(ps.lasthit, np.int64).tofile(o)
        np.ascontiguousarray(ps.hits, np.int32).tofile(o)
        np.ascontiguousarray(rays, psr.RAY_DTYPE).tofile(o)


def read_out(path, n):
    raw = open(path, "rb").read()
    out, pos = {}, 0
    for name, dt in [(x, np.float64) for x in F64] + [(x, np.float32) for x in ACC] + \
            [("ie", np.float32), ("lasthit", np.int64), ("hits", np.int32)]:
        size = np.dtype(dt).itemsize * n
        out[name] = np.frombuffer(raw[pos:pos + size], dt)
        pos += size
    out["total_deposit_count"] = int(np.frombuffer(raw[pos + 8:pos + 16], np.int64)[0])
    return out


def test_c_driver_is_built_against_the_header():
    assert os.access(DRIVER, os.X_OK), "run psphray_amd/build.sh"
    r = subprocess.run([DRIVER], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("cfg", [1, 3])
def test_c_driver_whole_tick_against_oracle(tmp_path, cfg):
    """workload 1 (hydrogen, isothermal) and 3 (H + He with temperature) at
    16^3 through the C caller: hits and last hits bit-exact, x_HI within
    1e-6 (isothermal) or 1e-3 (H + He), the deposit count equal"""
    print_case("c driver", cfg)
    wl = make_workload(cfg, RATES, n=16, nray=64)
    rays = wl.rays(1)
    case, outp = tmp_path / "case.bin", tmp_path / "out.bin"
    write_case(case, wl, rays, 1)
    r = subprocess.run([DRIVER, RATES, os.path.join(GOLDEN, "kernel.akline"), str(case), str(outp)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = read_out(outp, wl.ps.npar)
    u, f = wl.physics.units, wl.physics.flags
    of = po.OrFlags(f.isothermal, f.adiabatic, f.on_the_spot, f.h_only, f.secondary_ionization,
                    f.trace_only, f.no_photon, f.comoving, f.redshift, f.fake_temperature)
    orc = po.Oracle(po.units(hmf=u.C_HMF, hemf=u.C_HEMF), of, rates=RATES)
    ref = wl.ps.copy()
    t = po.Tree(ref.pos, ref.sml, wl.boxsize)
    rc, st = po.tick(orc, t, ref, rays, 1, wl.tick_time)
    assert rc == 0
    assert out["total_deposit_count"] == st.total_deposit_count > 0
    assert np.array_equal(out["hits"], ref.hits)
    assert np.array_equal(out["lasthit"], ref.lasthit)
    tol = 1e-6 if f.isothermal else 1e-3
    for n in F64:
        assert np.abs(out[n] - getattr(ref, n)).max() < tol, n
