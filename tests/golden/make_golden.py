"""Writes the golden fixtures under tests/golden/ (run in a container where
/root/reference and oracle/_ref exist; the fixtures are committed so the
tests run anywhere).

  atomic_rates_Hui.txt   data file of the reference's test directory
                         (test/atomic_rates_Hui.txt), read by the host as input
  kernel_akline.npy      the 4096 values of src/kernel.akline
  pluecker_kat.npz       random segment/box cases and the result of the
                         reference's own pluecker_ (src/pluecker-sse.c,
                         compiled by oracle/Makefile into oracle/_ref)
  sph_depth_kat.npz      r_h samples and the reference's sph_depth
  lsoda_robertson.npz    the reference LSODA on its Robertson demo
                         (support/lsoda/main.c) + the published values
  chem_lsoda.npz         chemistry steps integrated by the reference LSODA
                         with the oracle's right-hand side
"""
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
sys.path.insert(0, ROOT)

from oracle import pyoracle  # noqa: E402


def copy_data():
    for name in ("Hui",):
        shutil.copyfile(f"{REF}/test/atomic_rates_{name}.txt",
                        os.path.join(HERE, f"atomic_rates_{name}.txt"))
    vals = [float(l.strip().rstrip(",")) for l in open(f"{REF}/src/kernel.akline") if l.strip()]
    np.save(os.path.join(HERE, "kernel_akline.npy"), np.array(vals, np.float64))


def pluecker_kat(R, n=20000, seed=7):
    rng = np.random.default_rng(seed)
    dirs = rng.normal(size=(n, 3)).astype(np.float32)
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    # exact zeros in some components (class boundaries)
    z = rng.random((n, 3)) < 0.05
    dirs[z] = 0.0
    dirs = dirs.astype(np.float32)
    dist = rng.uniform(0, 4, n).astype(np.float32)
    # boxes centred near the segment (about half are hit), some with a face
    # placed exactly on a point of the segment (float rounding bands)
    t = rng.uniform(-0.5, 1.5, (n, 1)) * dist[:, None]
    ctr = t * dirs + rng.normal(0, 0.4, (n, 3))
    half = rng.uniform(0.01, 0.6, (n, 3))
    lo = (ctr - half).astype(np.float32)
    hi = (ctr + half).astype(np.float32)
    edge = rng.random(n) < 0.2
    ax = rng.integers(0, 3, n)
    pt = (t * dirs).astype(np.float32)
    for i in np.nonzero(edge)[0]:
        lo[i, ax[i]] = pt[i, ax[i]]
    s2b = lo.copy()
    s2t = hi.copy()
    res = np.zeros(n, np.int32)
    for i in range(n):
        d = np.ascontiguousarray(dirs[i])
        ds = np.array([dist[i]], np.float32)
        b = np.ascontiguousarray(s2b[i])
        t = np.ascontiguousarray(s2t[i])
        res[i] = R.pluecker_(d.ctypes.data, ds.ctypes.data, b.ctypes.data, t.ctypes.data)
    np.savez(os.path.join(HERE, "pluecker_kat.npz"), dir=dirs, dist=dist, s2b=s2b, s2t=s2t, hit=res)


def sph_depth_kat(R):
    r = np.concatenate([np.linspace(-0.01, 1.01, 5001), np.arange(4096) / 4096.0,
                        (np.arange(4096) + 0.999) / 4096.0])
    v = np.array([R.sph_depth(float(x)) for x in r])
    np.savez(os.path.join(HERE, "sph_depth_kat.npz"), r=r, v=v)


def robertson(R):
    out = np.zeros(36)
    rc = R.ref_robertson_run(out.ctypes.data)
    assert rc == 0
    # published LSODA demo answers (support/lsoda/main.c:47-60)
    pub = np.array([[9.851712e-01, 3.386380e-05, 1.479493e-02],
                    [9.055333e-01, 2.240655e-05, 9.444430e-02],
                    [7.158403e-01, 9.186334e-06, 2.841505e-01],
                    [4.505250e-01, 3.222964e-06, 5.494717e-01],
                    [1.831976e-01, 8.941773e-07, 8.168015e-01],
                    [3.898729e-02, 1.621940e-07, 9.610125e-01],
                    [4.936362e-03, 1.984221e-08, 9.950636e-01],
                    [5.161833e-04, 2.065787e-09, 9.994838e-01],
                    [5.179804e-05, 2.072027e-10, 9.999482e-01],
                    [5.283675e-06, 2.113481e-11, 9.999947e-01],
                    [4.658667e-07, 1.863468e-12, 9.999995e-01],
                    [1.431100e-08, 5.724404e-14, 1.000000e+00]])
    np.savez(os.path.join(HERE, "lsoda_robertson.npz"), ref=out.reshape(12, 3), published=pub)


def chem_steps(kind, n, rng, orc):
    """random Step inputs around the benchmark regimes"""
    u = orc.u
    s = np.zeros(n, pyoracle.STEP_DTYPE)
    nH = 10 ** rng.uniform(-4, -1, n) / u.U_CM ** 3
    s["nH"] = nH
    s["time"] = rng.choice([1, 2, 5, 20, 200], n) * 0.05 * u.U_MYR
    xhi = np.where(rng.random(n) < 0.5, 1 - 10 ** rng.uniform(-4, -0.3, n), 10 ** rng.uniform(-5, -0.3, n))
    s["lambdaH"] = xhi
    # deposits in units of the atoms (a fraction of a full ionization, sometimes several)
    s["yGdepHI"] = np.where(rng.random(n) < 0.3, 0.0, 10 ** rng.uniform(-4, 0.5, n))
    s["T"] = 1e4
    if kind == "h_iso":
        s["lambdaHeI"] = 1.0
        s["lambdaHeII"] = 0.0
    else:
        s["lambdaHeI"] = rng.uniform(0.0, 1.0, n)
        s["lambdaHeII"] = (1 - s["lambdaHeI"]) * rng.uniform(0, 1, n)
        s["yGdepHeI"] = np.where(rng.random(n) < 0.3, 0.0, 10 ** rng.uniform(-4, 0, n))
        s["yGdepHeII"] = np.where(rng.random(n) < 0.5, 0.0, 10 ** rng.uniform(-4, -1, n))
        T = 10 ** rng.uniform(3.5, 4.6, n)
        ye = (1 - s["lambdaH"]) + u.C_HEMF / u.C_HMF * 0.25 * (s["lambdaHeII"] + 2 * (1 - s["lambdaHeII"] - s["lambdaHeI"]))
        s["ie"] = [orc.Tye2ie(T[i], ye[i]) for i in range(n)]
        s["T"] = T
        # heat per mass comparable to a few eV per deposited photon
        s["heat"] = s["yGdepHI"] * u.C_H_PER_MASS * 3.0 * u.U_EV
    return s


def chem_lsoda():
    rng = np.random.default_rng(11)
    out = {}
    # isothermal H-only on-the-spot (workloads 1, 2, 4, 5)
    u = pyoracle.units(hmf=1.0)
    orc = pyoracle.Oracle(u, pyoracle.flags(isothermal=True, h_only=True, on_the_spot=True))
    s = chem_steps("h_iso", 400, rng, orc)
    o, st = orc.lsoda_evolve(s)
    out["h_iso_in"], out["h_iso_out"], out["h_iso_state"] = s, o, st
    # H+He with temperature (workload 3)
    u2 = pyoracle.units(hmf=0.76)
    orc2 = pyoracle.Oracle(u2, pyoracle.flags(isothermal=False, h_only=False, on_the_spot=True))
    s2 = chem_steps("he_temp", 300, rng, orc2)
    o2, st2 = orc2.lsoda_evolve(s2)
    out["he_in"], out["he_out"], out["he_state"] = s2, o2, st2
    np.savez(os.path.join(HERE, "chem_lsoda.npz"), **out)


if __name__ == "__main__":
    if not os.path.isdir(REF):
        sys.exit("needs /root/reference")
    copy_data()
    subprocess_ok = os.system(f"make -s -C {ROOT}/oracle all ref") == 0
    assert subprocess_ok
    R = pyoracle.ref_lib()
    pluecker_kat(R)
    sph_depth_kat(R)
    robertson(R)
    chem_lsoda()
    print("golden fixtures written")
