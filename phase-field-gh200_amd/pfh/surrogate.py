"""Learned surrogate of the 2D step: a BF16 CNN that advances every cell's
96x96 crop by K plain steps, trained on crops the plain solver generates,
with a plain check step per cell and a fallback to the plain method.

Network (kernels in csrc/sur_kernels.hip): input 8 channels (phi,
S_other = S - q(phi), cos theta, sin theta, 4 zero channels), three 3x3
convolutions of 32 channels with ReLU on MFMA (BF16 operands, FP32
accumulation), a 3x3 output convolution to one channel; the prediction is
phi + scale * out. Training: mean squared error on (phi_{n+K} - phi_n) /
scale, Adam on FP32 master weights, every gradient from the library's own
backward kernels.

Acceptance (the check-step rule of this build): for each cell, one explicit
plain step of size K dt from phi_n (the plain kernel with dt' = K dt, run on
copies of the state) gives a first-order reference of phi_{n+K}. Per-cell
rule (the default): a cell is accepted when, over its whole window, the
relative L2 distance E between the prediction and that reference is at most
`tol`, the distance relative to the block's own increment (reference -
phi_n) is at most `tol_inc` (a prediction of no change scores 1), and the
cell's accumulated committed error (the sum of E over its accepted blocks)
stays within `budget`. An accepted cell takes the prediction; a rejected
cell takes the reference field with the reference's window position and
class. The aggregate S, the moments and the K skipped polarity draws are
then recomputed exactly from the written fields (k_scatter_all /
k_init_moments / the plain tumble rule). Population rule
(fallback="population"): every cell must pass `tol` on its crop, otherwise
the whole block is advanced by K plain steps.
"""
from __future__ import annotations

import ctypes as C
import math
from pathlib import Path

import numpy as np
import torch

from . import abi

CROP = 96
PIX = CROP * CROP
C0 = 8
CH = 32
SIZES = {"W1": CH * 9 * C0, "b1": CH, "W2": CH * 9 * CH, "b2": CH, "W3": CH * 9 * CH,
         "b3": CH, "W4": 9 * CH, "b4": 1}
ORDER = list(SIZES)
NPARAM = sum(SIZES.values())
WEIGHTS_PATH = Path(__file__).resolve().parent / "data" / "surrogate_k8.npz"
# the same network trained for 16-step blocks (a variant beyond the 8-step
# check cadence; see pfh.surrogate_train)
WEIGHTS_PATH_K16 = WEIGHTS_PATH.with_name("surrogate_k16.npz")


def _ptr(t: torch.Tensor) -> int:
    return t.data_ptr()


class Net:
    """Device parameters: one fp32 master vector (views per tensor), bf16
    copies of the MFMA weights, Adam moments, the output scale."""

    def __init__(self, device="cuda:0", seed: int = 0, scale: float = 1e-3):
        self.dev = torch.device(device)
        self.lib = abi.load()
        self.w = torch.zeros(NPARAM, dtype=torch.float32, device=self.dev)
        self.m = torch.zeros_like(self.w)
        self.v = torch.zeros_like(self.w)
        self.g = torch.zeros_like(self.w)
        self.scale = float(scale)
        self.K = 8          # steps per block the weights were trained for
        self.t = 0
        self.views = {}
        self.gviews = {}
        o = 0
        for k in ORDER:
            self.views[k] = self.w[o:o + SIZES[k]]
            self.gviews[k] = self.g[o:o + SIZES[k]]
            o += SIZES[k]
        self.wb = {k: torch.zeros(SIZES[k], dtype=torch.int16, device=self.dev)
                   for k in ("W1", "W2", "W3")}
        self.wflip = {k: torch.zeros(SIZES[k], dtype=torch.int16, device=self.dev)
                      for k in ("W2", "W3")}
        self._init(seed)

    def _init(self, seed: int) -> None:
        rng = np.random.default_rng(seed)
        host = {}
        w1 = np.zeros((CH, 9, C0), np.float32)
        w1[:, :, :4] = rng.normal(0.0, math.sqrt(2.0 / 36.0), (CH, 9, 4))
        host["W1"] = w1
        host["W2"] = rng.normal(0.0, math.sqrt(2.0 / 288.0), (CH, 9, CH)).astype(np.float32)
        host["W3"] = rng.normal(0.0, math.sqrt(2.0 / 288.0), (CH, 9, CH)).astype(np.float32)
        host["W4"] = np.zeros((9, CH), np.float32)
        for k in ("b1", "b2", "b3", "b4"):
            host[k] = np.zeros(SIZES[k], np.float32)
        self.set_host(host)

    # ---- host transfer ----
    def set_host(self, host: dict) -> None:
        for k in ORDER:
            self.views[k].copy_(torch.from_numpy(np.ascontiguousarray(host[k], np.float32)
                                                 .reshape(-1)))
        self.refresh()

    def host(self) -> dict:
        return {k: self.views[k].cpu().numpy().copy() for k in ORDER}

    def refresh(self) -> None:
        """bf16 copies and the tap-flipped backward-data weights."""
        for k in ("W1", "W2", "W3"):
            self.wb[k].copy_(bf16_bits_torch(self.views[k]))
        for k in ("W2", "W3"):
            abi.check(self.lib.pfh_sur_flip(_ptr(self.wb[k]), _ptr(self.wflip[k]), None),
                      "pfh_sur_flip")

    def save(self, path: Path) -> None:
        path.parent.mkdir(parents=True, exist_ok=True)
        h = self.host()
        np.savez(path, scale=np.float64(self.scale), K=np.int64(self.K), **h)

    @classmethod
    def load(cls, path: Path = WEIGHTS_PATH, device="cuda:0") -> "Net":
        z = np.load(path, allow_pickle=False)
        net = cls(device=device, scale=float(z["scale"]))
        net.K = int(z["K"]) if "K" in z.files else 8
        net.set_host({k: z[k] for k in ORDER})
        return net

    def b4(self) -> float:
        return float(self.views["b4"].item())


def bf16_bits_torch(x: torch.Tensor) -> torch.Tensor:
    """binary32 -> bfloat16 bit patterns (round to nearest even) as int16."""
    u = x.contiguous().view(torch.int32).to(torch.int64) & 0xFFFFFFFF
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    return r.to(torch.int32).to(torch.int16)


class Workspace:
    """Activations and gradients for up to `n` crops."""

    def __init__(self, n: int, device, train: bool = False, hidden: bool = True):
        z = dict(device=device)
        self.n = n
        self.X0 = torch.zeros(n * PIX * C0, dtype=torch.int16, **z)
        # hidden maps of the layer-by-layer kernels (not needed by infer())
        self.h = [torch.zeros(n * PIX * CH, dtype=torch.int16, **z)
                  for _ in range(3 if hidden or train else 0)]
        self.phic = torch.zeros(n * PIX, dtype=torch.float32, **z)
        self.dphi = torch.zeros(n * PIX, dtype=torch.float32, **z)
        self.pred = torch.zeros(n * PIX, dtype=torch.float32, **z)
        self.origin = torch.zeros(2 * n, dtype=torch.int32, **z)
        if train:
            self.g = torch.zeros(n * PIX, dtype=torch.float32, **z)
            self.dh = [torch.zeros(n * PIX * CH, dtype=torch.int16, **z) for _ in range(3)]
            self.crop_loss = torch.zeros(n, dtype=torch.float64, **z)
            self.part = torch.zeros(n * (CH * 9 * CH + CH), dtype=torch.float32, **z)


def forward(net: Net, ws: Workspace, n: int, X0=None, phic=None, pred=True, stream=None):
    """dphi (and pred = phic + scale*out when pred) for n crops."""
    lib = net.lib
    X0 = ws.X0 if X0 is None else X0
    phic = ws.phic if phic is None else phic
    b = net.views
    abi.check(lib.pfh_sur_conv(8, _ptr(X0), _ptr(net.wb["W1"]), _ptr(b["b1"]), None,
                               _ptr(ws.h[0]), n, 1, stream), "conv1")
    abi.check(lib.pfh_sur_conv(32, _ptr(ws.h[0]), _ptr(net.wb["W2"]), _ptr(b["b2"]), None,
                               _ptr(ws.h[1]), n, 1, stream), "conv2")
    abi.check(lib.pfh_sur_conv(32, _ptr(ws.h[1]), _ptr(net.wb["W3"]), _ptr(b["b3"]), None,
                               _ptr(ws.h[2]), n, 1, stream), "conv3")
    abi.check(lib.pfh_sur_out(_ptr(ws.h[2]), _ptr(b["W4"]), net.b4(), _ptr(phic),
                              _ptr(ws.dphi), _ptr(ws.pred) if pred else None, n, stream),
              "out")


def infer(net: Net, ws: Workspace, n: int, W4=None, b4=None, dphi=True, stream=None, go=None):
    """The whole network in one kernel (pfh_sur_infer): ws.pred = ws.phic +
    dphi for n crops, and ws.dphi when `dphi`. W4/b4 default to the raw
    output layer (no scale)."""
    W4 = net.views["W4"] if W4 is None else W4
    b4 = net.b4() if b4 is None else b4
    b = net.views
    abi.check(net.lib.pfh_sur_infer(_ptr(ws.X0), _ptr(net.wb["W1"]), _ptr(b["b1"]),
                                    _ptr(net.wb["W2"]), _ptr(b["b2"]), _ptr(net.wb["W3"]),
                                    _ptr(b["b3"]), _ptr(W4), float(b4), _ptr(ws.phic),
                                    _ptr(ws.dphi) if dphi else None, _ptr(ws.pred), n,
                                    None if go is None else _ptr(go), stream),
              "pfh_sur_infer")


def infer_cells(net: Net, sim, ws: Workspace, W4=None, b4=None, dphi=True, stream=None,
                go=None):
    """infer() with the crops built inside the kernel from the Sim2D's field,
    aggregate and cell states (pfh_sur_infer_cells): ws.origin, ws.pred and,
    when `dphi`, ws.dphi; ws.X0 and ws.phic are not used."""
    W4 = net.views["W4"] if W4 is None else W4
    b4 = net.b4() if b4 is None else b4
    b = net.views
    s = sim.steps_done
    abi.check(net.lib.pfh_sur_infer_cells(
        _ptr(sim.phi[s % 2]), _ptr(sim.cell), _ptr(sim.cls), sim._s_ptr(s % 3), sim.N, sim.L,
        sim.P, _ptr(net.wb["W1"]), _ptr(b["b1"]), _ptr(net.wb["W2"]), _ptr(b["b2"]),
        _ptr(net.wb["W3"]), _ptr(b["b3"]), _ptr(W4), float(b4), _ptr(ws.origin),
        _ptr(ws.dphi) if dphi else None, _ptr(ws.pred), None if go is None else _ptr(go),
        sim._sp() if stream is None else stream), "pfh_sur_infer_cells")


def backward(net: Net, ws: Workspace, n: int, target: torch.Tensor, X0=None, stream=None):
    """Gradients into net.g for L = mean (dphi - target)^2 over n crops;
    returns the loss as a device scalar."""
    lib = net.lib
    X0 = ws.X0 if X0 is None else X0
    count = n * PIX
    abi.check(lib.pfh_sur_loss(_ptr(ws.dphi), _ptr(target), _ptr(ws.g), _ptr(ws.crop_loss), n,
                               2.0 / count, stream), "loss")
    gv = net.gviews
    # output layer: dh3 and dW4/db4 (partials laid out [n][289])
    abi.check(lib.pfh_sur_out_bwd(_ptr(ws.g), _ptr(ws.h[2]), _ptr(net.views["W4"]),
                                  _ptr(ws.dh[2]), _ptr(ws.part), n, stream), "out_bwd")
    tmp = torch.empty(9 * CH + 1, dtype=torch.float32, device=net.dev)
    abi.check(lib.pfh_sur_reduce(_ptr(ws.part), n, 9 * CH + 1, _ptr(tmp), stream), "reduce4")
    gv["W4"].copy_(tmp[:9 * CH])
    gv["b4"].copy_(tmp[9 * CH:])
    # hidden layers 3, 2, 1
    layers = [("W3", "b3", ws.h[1], 32), ("W2", "b2", ws.h[0], 32), ("W1", "b1", X0, 8)]
    dys = [ws.dh[2], ws.dh[1], ws.dh[0]]
    for i, (wk, bk, Xin, cin) in enumerate(layers):
        dY = dys[i]
        width = CH * 9 * cin + CH
        abi.check(lib.pfh_sur_wgrad(cin, _ptr(dY), _ptr(Xin), _ptr(ws.part), n, stream), "wgrad")
        red = torch.empty(width, dtype=torch.float32, device=net.dev)
        abi.check(lib.pfh_sur_reduce(_ptr(ws.part), n, width, _ptr(red), stream), "reduce")
        gv[wk].copy_(red[:CH * 9 * cin])
        gv[bk].copy_(red[CH * 9 * cin:])
        if cin == 32:
            # dX = conv(dY, flipped W) masked by the layer input's ReLU
            abi.check(lib.pfh_sur_conv(32, _ptr(dY), _ptr(net.wflip[wk]), None, _ptr(Xin),
                                       _ptr(dys[i + 1]), n, 0, stream), "dgrad")
    return ws.crop_loss[:n].sum() / count


def adam_step(net: Net, lr: float, b1=0.9, b2=0.999, eps=1e-8) -> None:
    net.t += 1
    c1 = 1.0 - b1 ** net.t
    c2 = 1.0 - b2 ** net.t
    abi.check(net.lib.pfh_sur_adam(_ptr(net.w), _ptr(net.g), _ptr(net.m), _ptr(net.v), None,
                                   NPARAM, lr, b1, b2, eps, c1, c2, None), "adam")
    net.refresh()


# ---------------------------------------------------------------------------
# data from the plain solver
# ---------------------------------------------------------------------------
def crop_sim(sim, ws: Workspace, out_X0=None, out_phic=None, out_origin=None) -> None:
    """Network inputs of every cell of a Sim2D at its current step."""
    lib = abi.load()
    s = sim.steps_done
    X0 = ws.X0 if out_X0 is None else out_X0
    phic = ws.phic if out_phic is None else out_phic
    origin = ws.origin if out_origin is None else out_origin
    abi.check(lib.pfh_sur_crop(_ptr(sim.phi[s % 2]), _ptr(sim.cell), _ptr(sim.cls),
                               sim._s_ptr(s % 3), sim.N, sim.L, sim.P, _ptr(X0), _ptr(phic),
                               _ptr(origin), sim._sp()), "pfh_sur_crop")


def recrop_sim(sim, origin: torch.Tensor, out: torch.Tensor) -> None:
    lib = abi.load()
    s = sim.steps_done
    abi.check(lib.pfh_sur_recrop(_ptr(sim.phi[s % 2]), _ptr(sim.cell), _ptr(sim.cls),
                                 _ptr(origin), sim.N, sim.L, _ptr(out), sim._sp()),
              "pfh_sur_recrop")


def make_dataset(seeds, samples_per_seed: int = 2, K: int = 8, warm: int = 300,
                 gap: int = 40, device="cuda:0", log=None):
    """Training crops from 2048^2 / 1024-cell monolayers (seeds as given):
    after `warm` plain steps, `samples_per_seed` samples `gap` steps apart;
    each sample is every cell's input crop at step n and the increment
    phi_{n+K} - phi_n on the same window."""
    from . import params, sim2d
    per = 1024 * samples_per_seed * len(seeds)
    dev = torch.device(device)
    X = torch.zeros(per * PIX * C0, dtype=torch.int16, device=dev)
    D = torch.zeros(per * PIX, dtype=torch.float32, device=dev)
    ws = Workspace(1024, dev)
    after = torch.zeros(1024 * PIX, dtype=torch.float32, device=dev)
    k = 0
    for seed in seeds:
        wl = params.monolayer(1024, 2048, 20.0, 32, 4.0, int(seed), name=f"train_{seed}")
        sim = sim2d.Sim2D(wl, device=device)
        sim.capture_graph()
        sim.run(warm)
        for s in range(samples_per_seed):
            if s:
                sim.run(gap)
            o = k * 1024
            crop_sim(sim, ws, out_X0=X[o * PIX * C0:(o + 1024) * PIX * C0])
            sim.run(K)
            recrop_sim(sim, ws.origin, after)
            D[o * PIX:(o + 1024) * PIX].copy_(after - ws.phic)
            k += 1
        if sim.fatal():
            raise RuntimeError(f"fatal flags in training run seed {seed}")
        del sim
        if log:
            log(f"dataset: seed {seed} done ({k * 1024} crops)")
    return X, D


def train(net: Net, X: torch.Tensor, D: torch.Tensor, epochs: int = 4, batch: int = 128,
          lr: float = 1e-3, seed: int = 0, log=None) -> list:
    """Adam over shuffled minibatches of (input, increment/scale)."""
    n = D.numel() // PIX
    ws = Workspace(batch, net.dev, train=True)
    tgt = torch.zeros(batch * PIX, dtype=torch.float32, device=net.dev)
    gen = torch.Generator(device="cpu").manual_seed(seed)
    hist = []
    Xv = X.view(n, PIX * C0)
    Dv = D.view(n, PIX)
    steps_total = epochs * (n // batch)
    step = 0
    for ep in range(epochs):
        perm = torch.randperm(n, generator=gen).to(net.dev)
        tot = 0.0
        for bi in range(n // batch):
            idx = perm[bi * batch:(bi + 1) * batch]
            ws.X0.view(batch, PIX * C0).copy_(Xv[idx])
            tgt.view(batch, PIX).copy_(Dv[idx] / net.scale)
            forward(net, ws, batch, pred=False)
            loss = backward(net, ws, batch, tgt)
            # cosine decay
            cur = lr * 0.5 * (1.0 + math.cos(math.pi * step / max(1, steps_total)))
            adam_step(net, cur)
            step += 1
            tot += float(loss)
        hist.append(tot / max(1, n // batch))
        if log:
            log(f"epoch {ep}: mean loss {hist[-1]:.4e}")
    return hist


def deploy_weights(net: Net) -> tuple[torch.Tensor, float]:
    """Output-layer weights with the scale folded in (fp32)."""
    return net.views["W4"] * net.scale, net.b4() * net.scale


# ---------------------------------------------------------------------------
# deployment on a Sim2D
# ---------------------------------------------------------------------------
def check_step_stable(p, K: int) -> bool:
    """Whether one explicit step of K dt is itself a stable integrator, so
    that a rejected cell may take it: the largest decay rate of the
    linearised update (the 9-point Laplacian's 32/6 gamma, the double well's
    bulk gamma at phi = 0 or 1, the interaction's rep S_other with
    S_other <= 1, each scaled by the mobility ratio) times K dt stays below
    2. At dt = 0.01 with the monolayer's coefficients this holds for K <= 11
    (K = 8: 1.46; K = 16: 2.9, where a fallback to the check field grows
    without bound)."""
    gam = max(p.gamma_normal, p.gamma_cancer if p.cancer_fraction > 0 else 0.0)
    mrel = p.phase_field_mobility / 0.5
    rate = mrel * (gam * 32.0 / 6.0 + p.bulk() * gam + p.interaction())
    return K * p.dt * rate < 2.0


class SurrogateRunner:
    """Advances a Sim2D K steps at a time with the surrogate, checked per
    cell against one explicit step of size K dt. Per-cell fallback (the
    default): a rejected cell takes that check step's field, so a block
    always advances. Population fallback: a block with any rejected cell is
    replaced by K plain steps.

    Blocks are issued in chunks with no host round trip inside a chunk. Every
    launch of a block is gated on a device flag `go`: propose (check-step
    copies, fused inference with in-kernel crops, the check step, the
    acceptance test into a rejected-cell count), then commit (polarity
    draws, exact S and moments, counters advanced by K), which does nothing
    when the count is non-zero. A rejection clears `go` and records the
    block's first step, so the rest of the chunk does nothing; after the
    chunk the host runs the K plain steps of that block, sets `go` again and
    issues the remaining blocks. With `graphs` the launches of a block are
    captured once per residue of the step index mod 6 (the phi parity and S
    rotation it touches) and replayed. In per-cell mode the rejected count
    stays zero and no chunk is ever interrupted."""

    def __init__(self, sim, net: Net, K: int | None = None, tol: float = 1e-3,
                 graphs: bool = True, chunk: int = 8, fallback: str | None = None,
                 tol_inc: float = 0.25, budget: float = 5e-4):
        K = net.K if K is None else int(K)
        if K != net.K:
            raise ValueError(f"the weights predict {net.K} steps; K = {K} would apply them "
                             f"as {K}")
        if K <= 0 or K % 2:
            raise ValueError("K must be a positive even number of steps")
        stable = check_step_stable(sim.p, K)
        if fallback is None:
            fallback = "cell" if stable else "population"
        if fallback not in ("cell", "population"):
            raise ValueError("fallback is 'cell' or 'population'")
        if fallback == "cell" and not stable:
            raise ValueError(f"a rejected cell would take one explicit step of {K} dt, which "
                             f"is not a stable integrator for these coefficients; use "
                             f"fallback='population'")
        self.sim, self.net, self.K, self.tol = sim, net, K, tol
        # per-cell fallback: a rejected cell takes the check step's field;
        # acceptance also bounds the error relative to the block's increment
        # and each cell's accumulated committed error (budget)
        self.fallback, self.tol_inc, self.budget = fallback, tol_inc, budget
        self.chunk = max(1, int(chunk))
        self.lib = abi.load()
        dev = sim.dev
        N = sim.N
        self.ws = Workspace(N, dev, hidden=False)
        self.W4d, self.b4d = deploy_weights(net)
        self.phi_chk = torch.zeros_like(sim.phi[0])
        self.cell_chk = torch.zeros_like(sim.cell)
        self.cls_chk = torch.zeros_like(sim.cls)
        self.cur_chk = torch.zeros(2, dtype=torch.int64, device=dev)
        self.step_chk = torch.zeros(2, dtype=torch.int64, device=dev)
        self.flags_chk = torch.zeros_like(sim.flags)
        self.err = torch.zeros(N, dtype=torch.float32, device=dev)
        # [0] go, [1] rejected cells of the current block
        self.ctl = torch.tensor([1, 0], dtype=torch.int32, device=dev)
        self.abort_step = torch.zeros(1, dtype=torch.int64, device=dev)
        # [0] largest check error so far (float bits), [1] rejected blocks
        # [2] cells that fell back (per-cell mode), summed over blocks
        self.counters = torch.zeros(3, dtype=torch.int32, device=dev)
        self.acc = torch.zeros(N, dtype=torch.float32, device=dev)
        self.accept = torch.zeros(N, dtype=torch.int32, device=dev)
        self.graphs = graphs
        self._g = {}
        self.blocks = 0

    def _check_args(self, slot: int) -> abi.StepArgs:
        sim = self.sim
        a = abi.StepArgs()
        C.memmove(C.addressof(a), C.addressof(sim._args[slot % 6]), C.sizeof(a))
        a.phi_out = _ptr(self.phi_chk)
        a.S_sc = None                  # the check step's aggregate is never read
        a.S_cl = None
        a.clear_ahead_words = 0
        a.cell = _ptr(self.cell_chk)
        a.cell_cls = _ptr(self.cls_chk)
        a.cursor_use = _ptr(self.cur_chk)
        a.cursor_clear = _ptr(self.cur_chk) + 8
        a.step_rd = _ptr(self.step_chk)
        a.step_wr = _ptr(self.step_chk) + 8
        a.flags = _ptr(self.flags_chk)
        a.dt = float(np.float32(sim.p.dt * self.K))
        a.physical_dt = sim.p.dt * self.K
        return a

    def _issue(self, s: int) -> None:
        """All launches of the block that starts at step s."""
        sim, lib, K = self.sim, self.lib, self.K
        sp = sim._sp()
        N = sim.N
        pin = s % 2
        e = s + K
        go = _ptr(self.ctl)
        rej = go + 4
        now = _ptr(sim.stepc) + 8 * pin
        # propose
        abi.check(lib.pfh_sur_prep(go, rej, _ptr(sim.cell), _ptr(self.cell_chk), _ptr(sim.cls),
                                   _ptr(self.cls_chk), N, None, 0, _ptr(self.cur_chk),
                                   _ptr(self.step_chk), now, sp), "prep")
        infer_cells(self.net, sim, self.ws, W4=self.W4d, b4=self.b4d, dphi=False, stream=sp,
                    go=self.ctl)
        a = self._check_args(s)
        abi.check(lib.pfh2d_launch_check_step_if(C.byref(a), go, sim.grid, sp), "check step")
        if self.fallback == "cell":
            # every cell is decided on its own; rejected cells take the check
            # step's field, so the block always advances (rej stays 0)
            abi.check(lib.pfh_sur_zero2(sim._s_ptr(e % 3), sim._s_ptr((e + 1) % 3), sim.words,
                                        None, sp), "zero2")
            abi.check(lib.pfh_sur_check_cells(
                _ptr(self.phi_chk), _ptr(self.cell_chk), _ptr(self.cls_chk), _ptr(sim.phi[pin]),
                _ptr(sim.cell), _ptr(sim.cls), _ptr(self.ws.origin), N, sim.L,
                _ptr(self.ws.pred), self.tol, self.tol_inc, self.budget, _ptr(self.acc),
                _ptr(self.err), _ptr(self.accept), _ptr(self.counters) + 8,
                _ptr(self.counters), sp), "check cells")
            abi.check(lib.pfh_sur_commit_cells(
                _ptr(sim.phi[pin]), _ptr(sim.cell), _ptr(sim.cls), _ptr(self.ws.origin),
                _ptr(self.ws.pred), _ptr(self.accept), _ptr(self.phi_chk), _ptr(self.cell_chk),
                _ptr(self.cls_chk), pin, N, sim.L, sim.P, sim._s_ptr(e % 3), _ptr(sim.flags),
                sp), "commit cells")
            abi.check(lib.pfh_sur_finish(_ptr(sim.cell), N, _ptr(self.step_chk), K,
                                         sim.p.polarity_stream(), sim.p.p_tumble(),
                                         _ptr(sim.stepc), _ptr(sim.cursor), pin, go, rej,
                                         _ptr(self.abort_step), _ptr(self.counters) + 4, sp),
                      "finish")
            return
        abi.check(lib.pfh_sur_check(_ptr(self.phi_chk), _ptr(self.cell_chk), _ptr(self.cls_chk),
                                    _ptr(self.ws.origin), N, sim.L, _ptr(self.ws.pred),
                                    self.tol, _ptr(self.err), rej, _ptr(self.counters), go, sp),
                  "check")
        # commit (nothing when the block has rejected cells): the next step
        # reads S[e%3] and scatters into S[(e+1)%3]; it clears S[(e+2)%3]
        # ahead by itself
        abi.check(lib.pfh_sur_zero2(sim._s_ptr(e % 3), sim._s_ptr((e + 1) % 3), sim.words, rej,
                                    sp), "zero2")
        abi.check(lib.pfh_sur_commit(_ptr(sim.phi[pin]), _ptr(sim.cell), _ptr(sim.cls),
                                     _ptr(self.ws.origin), _ptr(self.ws.pred), N, sim.L, sim.P,
                                     sim._s_ptr(e % 3), _ptr(sim.flags), rej, sp), "commit")
        # the K polarity draws (from the block's first step, which prep left
        # in step_chk[0]) and the counter advance, in one launch
        abi.check(lib.pfh_sur_finish(_ptr(sim.cell), N, _ptr(self.step_chk), K,
                                     sim.p.polarity_stream(), sim.p.p_tumble(), _ptr(sim.stepc),
                                     _ptr(sim.cursor), pin, go, rej, _ptr(self.abort_step),
                                     _ptr(self.counters) + 4, sp), "finish")

    def _captured(self, r: int) -> torch.cuda.CUDAGraph:
        g = self._g.get(r)
        if g is not None:
            return g
        sim = self.sim
        g = torch.cuda.CUDAGraph()
        st = torch.cuda.Stream(sim.dev)
        st.wait_stream(sim.stream)
        with torch.cuda.stream(st):
            old = sim.stream
            sim.stream = st
            try:
                with torch.cuda.graph(g, stream=st):
                    self._issue(r)
            finally:
                sim.stream = old
        sim.stream.wait_stream(st)
        torch.cuda.synchronize(sim.dev)
        self._g[r] = g
        return g

    def prepare(self) -> None:
        """Capture the graphs of the next three blocks' residues (capture
        runs nothing), so that timed blocks replay only."""
        if self.graphs:
            for b in range(3):
                self._captured((self.sim.steps_done + b * self.K) % 6)

    def _launch(self) -> None:
        sim = self.sim
        s = sim.steps_done
        if self.graphs:
            self._captured(s % 6).replay()
        else:
            self._issue(s)
        sim.steps_done = s + self.K

    def run(self, steps: int) -> None:
        """Advance `steps` (a multiple of K) steps."""
        sim, K = self.sim, self.K
        if steps % K:
            raise ValueError(f"steps must be a multiple of K = {K}")
        left = steps // K
        while left:
            m = min(self.chunk, left)
            s0 = sim.steps_done
            for _ in range(m):
                self._launch()
            self.blocks += m
            if int(self.ctl[0].item()):           # waits for the chunk
                left -= m
                continue
            # the block at abort_step was rejected and the blocks after it
            # did nothing: run its plain steps, then carry on from there
            sa = int(self.abort_step.item())
            done = (sa - s0) // K
            self.blocks -= m - done - 1
            sim.steps_done = sa
            sim.run(K)
            self.ctl.fill_(1)
            left -= done + 1

    def block(self) -> None:
        """Advance K steps (surrogate, or plain when rejected)."""
        self.run(self.K)

    def stats(self) -> dict:
        """Blocks run, rejected (fallback) blocks and the largest per-cell
        check error so far (waits for the stream)."""
        c = self.counters.cpu().numpy()
        out = {"blocks": self.blocks, "fallback_blocks": int(c[1]),
               "accepted_blocks": self.blocks - int(c[1]),
               "max_check_error": float(c[:1].view(np.float32)[0]), "fallback": self.fallback}
        if self.fallback == "cell":
            cells = self.blocks * self.sim.N
            out["fallback_cells"] = int(c[2])
            out["accepted_cell_fraction"] = 1.0 - int(c[2]) / max(1, cells)
            out["max_accumulated_error"] = float(self.acc.max().item())
        return out
