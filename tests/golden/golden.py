"""Golden fixtures of the FP64 oracle on tiny grids, and the script that
writes them (python tests/golden/golden.py).

The reference ships no tests, golden vectors or fixtures, so these vectors
come from the project's own FP64 oracle, written only after the oracle passed
its closed-form and invariant checks (tests/test_oracle3d.py,
tests/test_oracle2d.py, tests/test_cpml.py).  They pin the oracle against
drift (tests/test_golden.py re-runs it and compares) and give the GPU path a
fixed target that does not need the oracle at run time.

Cases (each one function of the case name, shared by the writer and the tests):
  3D model pass, lt steps, record + strided fields + every field's L2 norm:
    d4  damping boundary, radius 4      c4  C-PML, radius 4      c6  C-PML, radius 6
  3D excitation RTM of one shot (true-model record, smoothed migration): r4 images
  2D: e2 elastic forward (record, fields), Born data of a seeded dm and the
      adjoint of a seeded residual; v2 viscoelastic L = 3 forward.
Stored as float32 (fields strided by 2 in each axis in 3D) in
oracle_fixtures.npz, ~0.5 MB.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))

import cases  # noqa: E402
import cases2d  # noqa: E402

FIXTURES = os.path.join(HERE, "oracle_fixtures.npz")
FIELDS3D = ["vx", "vy", "vz", "txx", "tyy", "tzz", "txy", "txz", "tyz", "tp", "vxp", "vyp", "vzp"]
STRIDED3D = ["vx", "vz", "txx", "txy", "tp"]
FIELDS2D = ["vx", "vz", "txx", "tzz", "txz", "tp"]


def case3d(name):
    """(params, true model, smoothed model) of a 3D case."""
    R = 6 if name == "c6" else 4
    b = 10 if R == 6 else 8
    p = cases.params(nx=20, ny=18, nz=16, b=b, R=R, lt=40, dt=1.0, freq=25.0, src=(10, 9, 4))
    p.boundary = 0 if name in ("d4", "r4") else 1
    true = cases.layered(p, seed=5)
    sm = [cases.smooth(a, 2) for a in true]
    return p, true, sm


def case2d(name):
    """(params, model arrays, background arrays) of a 2D case."""
    L = 3 if name == "v2" else 0
    p = cases2d.params(nx=48, nz=36, b=10, R=4, L=L, lt=160, freq=25.0, src=(24, 3), rec_dx=1, rec_z=2)
    vp, vs, rho, qp, qs = cases2d.model(p, seed=3, q=L > 0)
    bg = [cases2d.smooth2(a) for a in (vp, vs, rho)] + ([qp, qs] if L > 0 else [None, None])
    return p, (vp, vs, rho, qp, qs), bg


def seeded_dm(p):
    rng = np.random.default_rng(11)
    dm = np.zeros((3, p.nx, p.nz), np.float32)
    dm[:2, 4:-4, 6:-4] = 0.05 * rng.standard_normal((2, p.nx - 8, p.nz - 10))
    return dm


def seeded_residual(p):
    return np.random.default_rng(12).standard_normal((2, p.lt, p.nrec)).astype(np.float32)


def outputs_3d(name, ctx_factory, model_pass=0):
    """The 3D fixture arrays of case `name` from an object with the
    Oracle3D / RTM3D interface (load_model, step, get_field, get_record)."""
    p, true, sm = case3d(name)
    out = {}
    if name == "r4":
        c = ctx_factory(p)
        c.load_model(0, *true)
        c.load_model(1, *sm)
        imgs = c.rtm_images()
        for k, a in zip(("img_tp", "img_pp", "img_ps"), imgs):
            out[f"{name}_{k}"] = np.asarray(a, np.float32)
        return out
    c = ctx_factory(p)
    c.load_model(0, *true)
    c.step(model_pass, 0, p.lt)
    out[f"{name}_record"] = np.asarray(c.get_record(), np.float32)
    norms = []
    for k in FIELDS3D:
        f = np.asarray(c.get_field(k), np.float64)
        norms.append(np.linalg.norm(f))
        if k in STRIDED3D:
            out[f"{name}_{k}"] = f[::2, ::2, ::2].astype(np.float32)
    out[f"{name}_norms"] = np.array(norms)
    return out


def outputs_2d(name, ctx_factory):
    p, model, bg = case2d(name)
    out = {}
    c = ctx_factory(p)
    c.load_model(0, *model)
    c.load_model(1, *bg)
    out[f"{name}_record"] = np.asarray(c.forward_record(), np.float32)
    for k in FIELDS2D:
        out[f"{name}_{k}"] = np.asarray(c.get_field(k), np.float32)
    if name == "e2":
        out[f"{name}_born"] = np.asarray(c.born(seeded_dm(p)), np.float32)
        out[f"{name}_adjoint"] = np.asarray(c.adjoint(seeded_residual(p)), np.float32)
    return out


class Oracle3DCase:
    """Oracle3D with the two helpers the fixture functions use."""

    def __init__(self, p):
        import oracle3d
        self.o = oracle3d.Oracle3D(p, 64)

    def __getattr__(self, k):
        return getattr(self.o, k)

    def rtm_images(self):
        self.o.rtm_shot(False)
        return self.o.images()


class Oracle2DCase:
    def __init__(self, p):
        import oracle2d
        self.p = p
        self.o = oracle2d.Oracle2D(p)

    def __getattr__(self, k):
        return getattr(self.o, k)

    def forward_record(self):
        self.o.zero(7)
        self.o.step(0, 0, self.p.lt)
        return self.o.record()


def oracle_outputs():
    out = {}
    for n in ("d4", "c4", "c6", "r4"):
        out.update(outputs_3d(n, Oracle3DCase))
    for n in ("e2", "v2"):
        out.update(outputs_2d(n, Oracle2DCase))
    return out


if __name__ == "__main__":
    out = oracle_outputs()
    np.savez_compressed(FIXTURES, **out)
    print(f"wrote {FIXTURES}: {len(out)} arrays, {os.path.getsize(FIXTURES) / 1e3:.0f} kB")
