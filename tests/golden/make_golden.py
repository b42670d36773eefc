"""Writes the golden fixtures under tests/golden/ by running the reference's
own translation units (oracle/_ref/vt_ref_run, built by `make -C oracle ref`
from /root/reference).  Each fixture is data only: the input parameters, the
reference state after N steps and derived numbers.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))

# name: (input-file keys, steps, rhs_at_end)
CASES = {
    # the shipped input with integration_order=8, full period T = 10*sqrt(2)
    "c1_p3_o8_8x8_full": (dict(p=3, nx=8, ny=8, integration_order=8), 10000, False),
    "c1_p3_o8_8x8_250": (dict(p=3, nx=8, ny=8, integration_order=8), 250, True),
    # the shipped input as is (integration_order=6)
    "shipped_p3_o6_8x8_250": (dict(p=3, nx=8, ny=8, integration_order=6), 250, True),
    "p2_o8_6x5_100": (dict(p=2, nx=6, ny=5, integration_order=8), 100, True),
    "p1_o4_7x6_100": (dict(p=1, nx=7, ny=6, integration_order=4), 100, True),
    "p0_o2_5x4_50": (dict(p=0, nx=5, ny=4, integration_order=2), 50, True),
    "euler_p3_o8_4x4_20": (dict(p=3, nx=4, ny=4, integration_order=8, stepping_method=0), 20, True),
    "lsq_p3_o8_4x4_20": (dict(p=3, nx=4, ny=4, integration_order=8, U_initialization_type=1), 20, True),
    "p3_o8_4x4_perturb1_20": (dict(p=3, nx=4, ny=4, integration_order=8, perturb_grid=1), 20, True),
}


def make(name, kw, steps, rhs):
    with tempfile.TemporaryDirectory() as d:
        r = O.run_reference(d, kw, steps, rhs_at_end=rhs)
    setup = r["setup"]
    out = dict(
        steps=np.int64(steps),
        p=np.int64(kw.get("p", 3)), nx=np.int64(kw.get("nx", 8)), ny=np.int64(kw.get("ny", 8)),
        integration_order=np.int64(kw.get("integration_order", 8)),
        stepping_method=np.int64(kw.get("stepping_method", 1)),
        init_type=np.int64(kw.get("U_initialization_type", 0)),
        perturb_grid=np.int64(kw.get("perturb_grid", 0)),
        dt=np.float64(float("14.14213562373") / 10000),
        type=setup[:, 0].astype(np.int32), neighbors=setup[:, 1:4].astype(np.int32),
        vertices=setup[:, 4:10].reshape(-1, 3, 2), geometry=setup[:, 10:],
        U0=r["U0"], F0=r["F0"], U=r["U"], F=r["F"], time=np.float64(r["time"]),
    )
    if rhs:
        out["dUdt"] = r["dUdt"]
    if out["p"] == 3:
        out["l2_state_error"] = O.l2_state_error(out["vertices"], r["U0"], r["U"])
        # as plottools/l2stateerror.py reads them: std::to_string text, 6 decimals
        txt = np.vectorize(lambda x: float("%f" % x))
        out["l2_state_error_text"] = O.l2_state_error(txt(out["vertices"]), txt(r["U0"]), txt(r["U"]))
    if len(out["type"]) > 64:  # keep fixtures small: F and geometry only for the small meshes
        for k in ("F0", "F", "geometry"):
            del out[k]
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **out)
    print(name, "written", r["stdout"].strip())


def flux_records(n=3000, seed=11):
    """Physical left/right states and unit normals: the vortex regime, wide
    ranges, and states with |u.n| near ci = c2^2 so that the entropy fix of
    every eigenvalue (Numericalflux.cpp:137-142) is exercised."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        kind = i % 3
        rec = []
        th = rng.uniform(0, 2 * np.pi)
        nv = np.array([np.cos(th), np.sin(th)])
        for side in range(2):
            rho = rng.uniform(0.99, 1.0) if kind == 0 else rng.uniform(0.3, 3.0)
            p = rng.uniform(2.8, 2.9) if kind == 0 else rng.uniform(0.2, 8.0)
            if kind == 2:
                c2 = 0.4 * (1.4 / 0.4 * p / rho)  # gmi * H estimate, ignoring kinetic terms
                speed = (c2 * c2) * rng.uniform(0.85, 1.15)
                vel = speed * nv * rng.choice([-1, 1]) + rng.uniform(-0.5, 0.5, 2)
            else:
                vel = rng.uniform(-0.3, 0.3, 2) + (0.7071 if kind == 0 else rng.uniform(-2, 2))
            E = p / (rho * 0.4) + 0.5 * (vel @ vel)
            rec += [rho, rho * vel[0], rho * vel[1], rho * E]
        out.append(rec + list(nv))
    return np.array(out)


def make_flux(name="roe_flux_known_answers"):
    import subprocess
    recs = flux_records()
    with tempfile.TemporaryDirectory() as d:
        fi, fo = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        recs.tofile(fi)
        subprocess.run([O.REF_RUN, "flux", fi, fo], check=True)
        flux = np.fromfile(fo).reshape(-1, 4)
    np.savez_compressed(os.path.join(OUT, name + ".npz"), records=recs, flux=flux)
    print(name, "written", len(recs))


# The reference's text output (grid/ and output/step_*/ files) of a small run,
# as its own writers produce it: input keys, steps.
TEXT_CASE = ("text_dumps_p3_o8_4x4_perturb1",
             dict(p=3, nx=4, ny=4, integration_order=8, perturb_grid=1, simulation_time="0.028284271247",
                  number_time_steps=20, write_every_steps=10), 20)


def make_text_dumps():
    import json
    name, kw, steps = TEXT_CASE
    files = {}
    with tempfile.TemporaryDirectory() as d:
        O.run_reference(d, kw, steps, write_dumps=True)
        for sub in ("grid", "output"):
            for dirpath, _, names in os.walk(os.path.join(d, sub)):
                for f in names:
                    path = os.path.join(dirpath, f)
                    with open(path, "rb") as fh:
                        files[os.path.relpath(path, d)] = fh.read().decode("ascii")
    with open(os.path.join(OUT, name + ".json"), "w") as fh:
        json.dump({"input": kw, "steps": steps, "files": dict(sorted(files.items()))}, fh, indent=0)
    print(name, "written", len(files), "files")


if __name__ == "__main__":
    names = sys.argv[1:] or list(CASES) + ["roe_flux_known_answers", TEXT_CASE[0]]
    for n in names:
        if n == TEXT_CASE[0]:
            make_text_dumps()
        elif n == "roe_flux_known_answers":
            make_flux()
        else:
            make(n, *CASES[n])
