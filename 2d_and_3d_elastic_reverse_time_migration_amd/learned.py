"""Learned image-domain preconditioner for 2D elastic LSRTM.

The workflow has no counterpart in the reference (which ships the LSRTM
operators only); every number it produces comes from the library.  The
approximate inverse Hessian is z = net(w * g): w is the diagonal from the
source illumination of the background model (erm2d_illumination,
w = 1 / (W + eps max W), mean 1) and the U-Net learns the rest.

1. random layered-plus-anomaly models (`random_model`), a smooth background
   and the relative perturbations (dvp/vp, dvs/vs) the Born operator is
   linearised in (erm2d.h, inversion_para 2 of the reference);
2. Born data of each model and CG-LSRTM on it (`erm2d_lsrtm_cg` with the
   diagonal w as preconditioner, all shots of a model batched in one launch
   per kernel); a CG monitor records, at selected iterations k, the gradient
   g_k = L^T(d - L m_k) and the model m_k.  The training pair is
   (w g_k, m_K - m_k): the image CG converges to after K iterations minus
   the current one, i.e. what the inverse Hessian would map the gradient to
   (target "converged", default; "true" uses dm - m_k instead);
3. 64x64 patches; the input is divided by its RMS and each target by its
   own RMS (the net learns the direction of the update, CG's line search its
   length, so early and late iterations, whose gradient-to-update ratios
   differ by orders of magnitude, train together); the U-Net trains on the
   GPU (erm_unet_train: BF16-MFMA forward/backward, Adam);
4. on a held-out model: plain CG, CG with the diagonal alone, and flexible
   CG whose iterations alternate z = net(w g) and z = w g (schedule "alt2",
   default), plus learned warm starts (k net iterations, then the diagonal):
   iterations and seconds to a data misfit |L m - d| / |d| <= tol (the
   illumination's cost included), and model errors against the plain
   converged image and the true perturbation.

The test suite runs a reduced version of the same steps
(tests/test_gpu_learned.py).
"""
import json
import time

import numpy as np

from . import DiagPrecond, Elastic2D, UNet
from .params2d import Params2D


def _smooth(a, sigma):
    """Separable Gaussian smoothing with edge replication (FP64)."""
    a = np.asarray(a, np.float64)
    r = max(1, int(3 * sigma))
    k = np.exp(-0.5 * (np.arange(-r, r + 1) / sigma) ** 2)
    k /= k.sum()
    for ax in (0, 1):
        pad = [(0, 0), (0, 0)]
        pad[ax] = (r, r)
        ap = np.pad(a, pad, mode="edge")
        a = np.apply_along_axis(lambda v: np.convolve(v, k, "valid"), ax, ap)
    return a


def random_model(nx, nz, seed, smooth_sigma=6.0, top=6):
    """Layered model with curved interfaces and Gaussian anomalies.

    Returns (background vp, vs, rho) as float32 [nx][nz] (z fastest) and the
    relative perturbation dm [3][nx][nz] (dvp/vp, dvs/vs, 0) of the true
    model against the background; dm is tapered to zero on the top `top`
    cells (sources and receivers) and on the side and bottom 4 cells."""
    rng = np.random.default_rng(seed)
    x = np.arange(nx)[:, None].astype(np.float64)
    z = np.arange(nz)[None, :].astype(np.float64)
    nl = int(rng.integers(3, 7))
    depths = np.sort(rng.uniform(0.12, 0.95, nl)) * nz
    vel = 1900.0 + np.cumsum(rng.uniform(80.0, 450.0, nl + 1))
    vp = np.full((nx, nz), vel[0])
    for i, d0 in enumerate(depths):
        slope = rng.uniform(-0.15, 0.15)
        amp = rng.uniform(0.0, 0.06) * nz
        lam = rng.uniform(0.5, 2.0) * nx
        phase = rng.uniform(0, 2 * np.pi)
        zi = d0 + slope * (x - nx / 2) + amp * np.sin(2 * np.pi * x / lam + phase)
        vp = np.where(z >= zi, vel[i + 1], vp)
    for _ in range(int(rng.integers(1, 4))):
        cx, cz = rng.uniform(0.1, 0.9) * nx, rng.uniform(0.25, 0.9) * nz
        s = rng.uniform(3.0, 10.0)
        vp += rng.uniform(-350.0, 350.0) * np.exp(-((x - cx) ** 2 + (z - cz) ** 2) / (2 * s * s))
    n = _smooth(rng.standard_normal((nx, nz)), 8.0)
    ratio = np.clip(1.732 * (1.0 + 0.05 * n / n.std()), 1.55, 2.0)
    vs = vp / ratio
    bvp, bvs = _smooth(vp, smooth_sigma), _smooth(vs, smooth_sigma)
    brho = 230.0 * bvp ** 0.25
    w = np.zeros((nx, nz))
    w[4:nx - 4, top:nz - 4] = 1.0
    w = _smooth(w, 2.0) * (w > 0)
    dm = np.zeros((3, nx, nz))
    dm[0] = (vp / bvp - 1.0) * w
    dm[1] = (vs / bvs - 1.0) * w
    f = lambda a: np.ascontiguousarray(a, np.float32)
    return f(bvp), f(bvs), f(brho), f(dm)


def survey_params(nx, nz, lt, freq=15.0, d=10.0, dt=1.0, b=20, fd_order=8):
    p = Params2D()
    p.nx, p.nz, p.dx, p.dz, p.dt, p.freq, p.lt = nx, nz, d, d, dt, freq, lt
    p.lt_forward = lt
    p.bu = p.bd = p.bl = p.br = b
    p.coe_attenuation = 200.0
    p.fd_order = fd_order
    p.nsls = 0
    p.shot_x, p.shot_z = nx // 2, 2
    p.nrec, p.rec_x0, p.rec_dx, p.rec_z = nx - 2, 1, 1, 2
    return p


def shot_positions(nx, nshots):
    xs = np.linspace(4, nx - 5, nshots).round().astype(int)
    return [(int(v), 2) for v in xs]


def born_data(g, shots, dm):
    out = []
    for sx, sz in shots:
        g.set_shot(sx, sz)
        out.append(g.born(dm))
    return np.stack(out)


def patches(img, size=64, stride=32):
    """[c][nx][nz] -> [n][c][size][size] with the given stride (both axes)."""
    c, nx, nz = img.shape
    xs = list(range(0, max(nx - size, 0) + 1, stride))
    zs = list(range(0, max(nz - size, 0) + 1, stride))
    return np.stack([img[:, i:i + size, k:k + size] for i in xs for k in zs])


def illum_weights(g, shots, eps=1e-3):
    """Diagonal preconditioner weights from the source illumination W of the
    loaded background model (erm2d_illumination): w = 1 / (W + eps max W),
    normalised to mean 1; returns (w, seconds)."""
    t0 = time.perf_counter()
    W = g.illumination(shots).astype(np.float64)
    w = 1.0 / (W + eps * W.max())
    return (w / w.mean()).astype(np.float32), time.perf_counter() - t0


def training_pairs(g, shots, dm, iters, keep, pmask=3, w=None, target="true"):
    """CG on Born data of dm (plain, or preconditioned by the diagonal w);
    returns the list of (x_k, t_k) for k in keep, x_k = g_k (plain) or
    w * g_k (the net's input in the weighted preconditioner), 2-channel
    images, and the misfit history.  target "true": t_k = dm - m_k, the
    update to the model the data were made from; "converged": t_k =
    m_K - m_k with m_K the LSRTM image after all `iters` iterations, the
    update CG itself was converging to (an inverse-Hessian image pair)."""
    d = born_data(g, shots, dm)
    rec = {}

    def mon(it, m, grad, mis):
        if it in keep:
            rec[it] = (grad[:2].copy(), m[:2].copy())

    pre = None if w is None else DiagPrecond(w)
    m, mis, sec, it = g.lsrtm_cg(shots, d, iters, pmask=pmask, monitor=mon, precond=pre)
    scale = 1.0 if w is None else w[None]
    goal = dm[:2] if target == "true" else m[:2]
    pairs = [((scale * rec[k][0]).astype(np.float32), (goal - rec[k][1]).astype(np.float32)) for k in sorted(rec)]
    return pairs, mis


def normalise_pairs(pairs, scale=None, per_pair=False):
    """x = g / rms(g).  Global: t = scale * target / rms(g) with one scale
    (the median over pairs of rms(g) / rms(target), unless given).  Per
    pair: t = target / rms(target), so the net learns the direction of the
    update only (CG's line search sets its length), which keeps pairs from
    late iterations, whose gradient-to-update ratio differs by orders of
    magnitude, on one footing."""
    rg = np.array([np.sqrt(np.mean(p[0].astype(np.float64) ** 2)) for p in pairs])
    rt = np.array([np.sqrt(np.mean(p[1].astype(np.float64) ** 2)) for p in pairs])
    xs = [(p[0] / r).astype(np.float32) for p, r in zip(pairs, rg)]
    if per_pair:
        return xs, [(p[1] / max(r, 1e-30)).astype(np.float32) for p, r in zip(pairs, rt)], None
    if scale is None:
        scale = float(np.median(rg / np.maximum(rt, 1e-30)))
    ts = [(scale * p[1] / r).astype(np.float32) for p, r in zip(pairs, rg)]
    return xs, ts, scale


def make_patch_set(xs, ts, size=64, stride=32, flip=True):
    X, T = [], []
    for x, t in zip(xs, ts):
        px, pt = patches(x, size, stride), patches(t, size, stride)
        X.append(px)
        T.append(pt)
        if flip:  # mirror in x: the survey (evenly spaced shots, full receiver line) is symmetric
            X.append(px[:, :, ::-1, :])
            T.append(pt[:, :, ::-1, :])
    return np.ascontiguousarray(np.concatenate(X), np.float32), np.ascontiguousarray(np.concatenate(T), np.float32)


class Scheduled:
    """Preconditioner that applies the net, z = net(w g), at the CG
    iterations `use(k)` selects (k = 0, 1, ... counts applications) and the
    diagonal, z = w g, at the others; both are positively homogeneous, so
    the flexible CG may mix them.  Calls the C adaptor erm_unet_precond."""

    def __init__(self, net, w, shape, use):
        from . import lib, _fptr
        self._lib, self._fptr = lib, _fptr
        self.net, self.w, self.use, self.k = net, w, use, 0
        nx, nz = shape
        if lib.erm_unet_bind_model(net._h, nx, nz) or lib.erm_unet_bind_weights(net._h, _fptr(w)):
            raise RuntimeError("binding the net to the model failed")

    def __call__(self, g):
        k, self.k = self.k, self.k + 1
        if not self.use(k):
            return (self.w[None] * g).astype(np.float32)
        z = np.empty_like(g)
        rc = self._lib.erm_unet_precond(self._fptr(np.ascontiguousarray(g, np.float32)), self._fptr(z), self.net._h)
        if rc:
            raise RuntimeError(f"erm_unet_precond returned {rc}")
        return z


SCHEDULES = {
    "alt2": lambda k: k % 2 == 0,          # net, diagonal, net, ...
    "alt3": lambda k: k % 3 == 0,          # net every third iteration
    "first3_alt2": lambda k: k < 3 or k % 2 == 0,
}


def time_to(mis, sec, tol):
    """(iterations, seconds) at which the misfit first reaches tol, or (None, None)."""
    idx = np.nonzero(np.asarray(mis) <= tol)[0]
    if len(idx) == 0:
        return None, None
    return int(idx[0]), float(sec[idx[0]])


KEEP = (0, 1, 2, 3, 5, 8, 12, 17, 23)  # CG iterations whose pairs train the net (of CG_ITERS)
CG_ITERS = 30


def generate(train_models=160, train_shape=(128, 96), train_shots=6, train_lt=1100, cg_iters=CG_ITERS,
             keep=KEEP, seed0=1000, illum_eps=1e-3, target="converged", log=print):
    """Training pairs from CG on random models (seeds seed0, seed0+1, ...):
    plain CG, or with illum_eps set, CG preconditioned by the illumination
    diagonal (pairs (w g_k, t_k), t_k as in training_pairs).  All shots of
    a model run batched in one launch per kernel (erm2d_set_batch)."""
    nx, nz = train_shape
    g = Elastic2D(survey_params(nx, nz, train_lt))
    keep = set(keep)
    pairs = []
    shots = shot_positions(nx, train_shots)
    for i in range(train_models):
        bvp, bvs, brho, dm = random_model(nx, nz, seed0 + i)
        g.load_model(1, bvp, bvs, brho)
        if i == 0:
            g.set_batch(len(shots))
        w = None if illum_eps is None else illum_weights(g, shots, illum_eps)[0]
        pr, mis = training_pairs(g, shots, dm, cg_iters, keep, w=w, target=target)
        pairs += pr
        log(f"train model {i}: {'plain' if w is None else 'illumination-preconditioned'} CG misfit"
            f" {mis[0]:.3f} -> {mis[-1]:.4f} ({len(pr)} pairs)")
    g.close()
    return pairs


def train(pairs, channels=(16, 32, 64, 64), epochs=60, batch=32, lr=1e-3, seed=0, per_pair=True, log=print):
    """Normalise, cut patches and train a fresh U-Net; returns (net, stats)."""
    xs, ts, scale = normalise_pairs(pairs, per_pair=per_pair)
    X, T = make_patch_set(xs, ts)
    net = UNet(channels, seed=seed)
    t0 = time.perf_counter()
    losses = net.train(X, T, batch=batch, epochs=epochs, lr=lr, seed=seed)
    st = {"patches": int(X.shape[0]), "target_scale": scale, "train_s": time.perf_counter() - t0}
    ms, flops = net.last_stats()
    st["train_device_ms"] = ms
    st["train_mfma_tflops"] = flops / ms / 1e9 if ms > 0 else None
    st["loss_first_last"] = [float(losses[0]), float(losses[-1])]
    st["loss_identity"] = float(np.mean((X - T) ** 2))  # loss of the identity map, for scale
    log(f"trained {X.shape[0]} patches, loss {losses[0]:.4f} -> {losses[-1]:.4f} (identity {st['loss_identity']:.4f}),"
        f" {st['train_s']:.1f} s, {st['train_mfma_tflops']:.2f} TFLOP/s")
    return net, st


def evaluate(net, test_shape=(256, 96), test_shots=12, test_lt=1100, test_iters=400, tol=1e-2, batch_shots=0,
             test_seed=5000, log=print, every=25, warm=5, flexible=False, illum_eps=1e-3, schedules=("alt2",),
             plain_file=None, plain_only=False):
    """Plain CG against CG preconditioned by `net` on a held-out model.

    With illum_eps set, the approximate inverse Hessian is z = net(w g) with
    the illumination diagonal w (the net was trained on w g); the diagonal
    alone (z = w g) is run as well, so the net's own contribution is
    visible, and the illumination's cost is included in the times of both.
    Warm starts run `warm` learned iterations and then continue with the
    net-free preconditioner (plain CG, or the diagonal).  For runs longer
    than one sitting: plain_only runs plain CG alone and saves it to
    plain_file (.npz); with plain_file given otherwise, plain CG is read from
    it instead of being run (same model, shots and data: all seeded)."""
    nx, nz = test_shape
    g = Elastic2D(survey_params(nx, nz, test_lt))
    bvp, bvs, brho, dm = random_model(nx, nz, test_seed)
    g.load_model(1, bvp, bvs, brho)
    shots = shot_positions(nx, test_shots)
    g.set_batch(len(shots) if batch_shots <= 0 else min(len(shots), batch_shots))
    d = born_data(g, shots, dm)

    def progress(name):
        def mon(it, m, grad, mis):
            if it % every == 0:
                log(f"{name} CG iteration {it}: misfit {mis:.4f}")
        return mon

    res = {}
    nrm = lambda a: float(np.linalg.norm(a[:2] - dm[:2]) / np.linalg.norm(dm[:2]))
    if plain_file and not plain_only:
        z = np.load(plain_file, allow_pickle=False)
        m_plain, mis_p, sec_p = z["m"], z["misfit"], z["seconds"]
        res["plain_from"] = plain_file
    else:
        m_plain, mis_p, sec_p, _ = g.lsrtm_cg(shots, d, test_iters, pmask=3, tol=tol, monitor=progress("plain"))
        if plain_file:
            np.savez(plain_file, m=m_plain, misfit=mis_p, seconds=sec_p)
    ip, tp = time_to(mis_p, sec_p, tol)
    res["plain"] = dict(misfit=[float(v) for v in mis_p], seconds=[float(v) for v in sec_p], iters_to_tol=ip,
                        s_to_tol=tp)
    res["model_err_vs_true"] = {"plain": nrm(m_plain)}
    if plain_only:
        g.close()
        return res
    w, t_w, base = None, 0.0, None
    if illum_eps is not None:
        w, t_w = illum_weights(g, shots, illum_eps)
        base = DiagPrecond(w)
        res["illumination_s"] = t_w
        m_d, mis_d, sec_d, _ = g.lsrtm_cg(shots, d, test_iters, pmask=3, tol=tol, precond=base,
                                          monitor=progress("illumination"))
        sec_d = sec_d + t_w
        i_d, t_d = time_to(mis_d, sec_d, tol)
        res["illumination"] = dict(eps=illum_eps, misfit=[float(v) for v in mis_d], seconds=[float(v) for v in sec_d],
                                   iters_to_tol=i_d, s_to_tol=t_d)
        res["model_err_vs_true"]["illumination"] = nrm(m_d)
        if tp and t_d:
            res["time_ratio_illumination"] = t_d / tp
        log(f"held-out: plain {ip} it / {tp} s, illumination diagonal {i_d} it / {t_d} s to misfit {tol}")
    if flexible:
        m_learn, mis_l, sec_l, _ = g.lsrtm_cg(shots, d, test_iters, pmask=3, tol=tol, precond=net, weights=w,
                                              monitor=progress("learned"))
        sec_l = sec_l + t_w
        il, tl = time_to(mis_l, sec_l, tol)
        res["learned"] = dict(misfit=[float(v) for v in mis_l], seconds=[float(v) for v in sec_l], iters_to_tol=il,
                              s_to_tol=tl, model_err_vs_plain=float(np.linalg.norm(m_learn[:2] - m_plain[:2])
                                                                    / np.linalg.norm(m_plain[:2])))
        res["model_err_vs_true"]["learned"] = nrm(m_learn)
        if tp and tl:
            res["time_ratio"] = tl / tp
        log(f"held-out: plain {ip} it / {tp} s, learned {il} it / {tl} s to misfit {tol}")
    for name in schedules:
        if w is None:
            break
        pre = Scheduled(net, w, (nx, nz), SCHEDULES[name])
        m_s, mis_s, sec_s, _ = g.lsrtm_cg(shots, d, test_iters, pmask=3, tol=tol, precond=pre,
                                          monitor=progress(name))
        sec_s = sec_s + t_w
        i_s, t_s = time_to(mis_s, sec_s, tol)
        key = "schedule_" + name
        res[key] = dict(misfit=[float(v) for v in mis_s], seconds=[float(v) for v in sec_s], iters_to_tol=i_s,
                        s_to_tol=t_s, model_err_vs_plain=float(np.linalg.norm(m_s[:2] - m_plain[:2])
                                                               / np.linalg.norm(m_plain[:2])))
        res["model_err_vs_true"][key] = nrm(m_s)
        log(f"held-out: plain {ip} it / {tp} s, schedule {name} {i_s} it / {t_s} s to misfit {tol}")
    warms = [warm] if isinstance(warm, int) else list(warm or [])
    for wk in [v for v in warms if v]:
        m_w, mis_w1, sec_w1, it_w = g.lsrtm_cg(shots, d, wk, pmask=3, tol=tol, precond=net, weights=w)
        m_w, mis_w2, sec_w2, _ = g.lsrtm_cg(shots, d, test_iters - it_w, pmask=3, m0=m_w, tol=tol, precond=base,
                                            monitor=progress(f"warm-start {wk}"))
        mis_w = np.concatenate([mis_w1, mis_w2[1:]])
        sec_w = np.concatenate([sec_w1, sec_w1[-1] + sec_w2[1:]]) + t_w
        iw, tw = time_to(mis_w, sec_w, tol)
        key = "warm_start" if len(warms) == 1 else f"warm_start_{wk}"
        res[key] = dict(warm_iters=wk, misfit=[float(v) for v in mis_w], seconds=[float(v) for v in sec_w],
                        iters_to_tol=iw, s_to_tol=tw, model_err_vs_plain=float(np.linalg.norm(m_w[:2] - m_plain[:2])
                                                                               / np.linalg.norm(m_plain[:2])))
        res["model_err_vs_true"][key] = nrm(m_w)
        if tp and tw:
            res["time_ratio_" + key] = tw / tp
        log(f"held-out: plain {ip} it / {tp} s, warm start ({wk} learned iterations) {iw} it / {tw} s"
            f" to misfit {tol}")
    g.close()
    return res


def run(train_models=160, train_shape=(128, 96), train_shots=6, train_lt=1100, cg_iters=CG_ITERS, keep=KEEP,
        test_shape=(256, 96),
        test_shots=12, test_lt=1100, test_iters=400, tol=1e-2, epochs=20, batch=32, lr=1e-3,
        channels=(16, 32, 64, 64), batch_shots=0, seed0=1000, test_seed=5000, warm=3, flexible=False, illum_eps=1e-3,
        target="converged", per_pair=True, schedules=("alt2",), log=print, params_in=None, params_out=None,
        plain_file=None, plain_only=False):
    """The whole workflow; returns a dict of results.  params_out: save the
    trained net's parameters (numpy .npy); params_in: skip data generation
    and training and evaluate a net with these saved parameters."""
    res = {"config": dict(train_models=train_models, train_shape=train_shape, train_shots=train_shots,
                          train_lt=train_lt, cg_iters=cg_iters, keep=list(keep), test_shape=test_shape,
                          test_shots=test_shots,
                          test_lt=test_lt, tol=tol, epochs=epochs, batch=batch, lr=lr, channels=channels,
                          illum_eps=illum_eps, target=target, per_pair=per_pair)}
    if plain_only:
        net = None
    elif params_in:
        net = UNet(channels, seed=seed0)
        net.set_params(np.load(params_in, allow_pickle=False))
        res["params_in"] = params_in
    else:
        t0 = time.perf_counter()
        pairs = generate(train_models, train_shape, train_shots, train_lt, cg_iters, keep=keep, seed0=seed0,
                         illum_eps=illum_eps, target=target, log=log)
        res["datagen_s"] = time.perf_counter() - t0
        net, st = train(pairs, channels, epochs, batch, lr, seed=seed0, per_pair=per_pair, log=log)
        res.update(st)
        if params_out:
            np.save(params_out, net.get_params())
    if test_iters <= 0:
        if net:
            net.close()
        return res
    res.update(evaluate(net, test_shape, test_shots, test_lt, test_iters, tol, batch_shots, test_seed, log=log,
                        warm=warm, flexible=flexible, illum_eps=illum_eps, schedules=schedules,
                        plain_file=plain_file, plain_only=plain_only))
    if net:
        net.close()
    return res


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="learned-preconditioner LSRTM workflow")
    ap.add_argument("--out", default=None)
    ap.add_argument("--train-models", type=int, default=160)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--test-iters", type=int, default=400)
    ap.add_argument("--batch-shots", type=int, default=0, help="shots per batched launch (0: all shots of a model)")
    ap.add_argument("--warm", type=int, nargs="*", default=[3],
                    help="learned warm-start lengths to compare (net iterations before the net-free CG)")
    ap.add_argument("--flexible", action="store_true", help="also run CG with the net at every iteration")
    ap.add_argument("--schedules", nargs="*", default=["alt2"], choices=sorted(SCHEDULES),
                    help="net/diagonal schedules of the flexible CG to compare")
    ap.add_argument("--cg-iters", type=int, default=CG_ITERS, help="CG iterations per training model")
    ap.add_argument("--keep", type=int, nargs="*", default=list(KEEP),
                    help="CG iterations whose (gradient, remaining update) pairs are kept for training")
    ap.add_argument("--target", choices=["true", "converged"], default="converged",
                    help="training target: dm - m_k (true) or m_K - m_k with m_K the image after --cg-iters")
    ap.add_argument("--global-scale", action="store_true",
                    help="one global target scale instead of normalising each target by its own RMS")
    ap.add_argument("--illum-eps", type=float, default=1e-3,
                    help="illumination diagonal w = 1/(W + eps max W) composed with the net (train and test)")
    ap.add_argument("--no-illum", action="store_true", help="no diagonal: the net alone on plain-CG gradients")
    a = ap.parse_args(argv)
    if any(k >= a.cg_iters or k < 0 for k in a.keep):
        ap.error("--keep iterations must lie in [0, --cg-iters)")
    res = run(train_models=a.train_models, cg_iters=a.cg_iters, keep=a.keep, epochs=a.epochs,
              test_iters=a.test_iters, batch_shots=a.batch_shots,
              warm=a.warm, flexible=a.flexible, illum_eps=None if a.no_illum else a.illum_eps,
              target=a.target, per_pair=not a.global_scale, schedules=tuple(a.schedules),
              log=lambda s: print(s, flush=True))
    txt = json.dumps(res)
    if a.out:
        with open(a.out, "w") as f:
            f.write(txt)
    print(txt)


if __name__ == "__main__":
    main()
