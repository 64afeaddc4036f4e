"""The differentiable scalar step (mx_adv_scal_step) and its hand-written
adjoint (mx_adv_scal_step_vjp) on a periodic box with a discretely
divergence-free MAC velocity:

* the GPU step equals the oracle's step bit for bit, in the advective and
  the conservative form of MakeEdgeScal;
* dot-product test: <v, J w> from the oracle's forward-mode derivative
  (dual numbers, oracle/oracle_step.cpp) equals <J^T v, w> from the GPU
  adjoint, for random w in the state and the velocity, to 1e-12 relative
  (ppm_type 1 and 2);
* the torch.autograd.Function of one step returns the adjoint's gradients;
* checkpointing every 5 steps gives the same gradient as no checkpoint;
* central finite differences (eps = 1e-6) of a 20-step rollout's loss at
  64 random cells of the initial state and 64 random faces of the velocity
  agree with the adjoint gradient to 1e-6 relative.  A point whose central
  differences with eps and eps/2 disagree (a limiter switch within eps) is
  reported and left out; at most a few may be.
"""
import numpy as np
import pytest

import pyoracle
from maestroex_amd.adjoint import periodic_mac_velocity

pytestmark = pytest.mark.gpu


def setup_case(n=16, ncomp=2, seed=0):
    rng = np.random.default_rng(seed)
    u = periodic_mac_velocity(n, seed=seed + 1)
    x = (np.arange(n) + 0.5) / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    s = np.stack([np.exp(-((X - 0.5) ** 2 + (Y - 0.4 - 0.1 * c) ** 2 + (Z - 0.5) ** 2) / 0.02)
                  + 0.05 * rng.standard_normal((n, n, n)) for c in range(ncomp)])
    dt = 0.5 / n
    return s, u, dt, 1e-8


def stepper_for(n, ncomp, dt, rel_eps, ppm_type=1):
    from maestroex_amd.adjoint import ScalStepper
    return ScalStepper(n, ncomp, dt, rel_eps, ppm_type=ppm_type)


def to_dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


@pytest.mark.parametrize("ppm", [1, 2])
def test_step_forward_parity(ppm):
    s, u, dt, eps = setup_case()
    n, nc = s.shape[1], s.shape[0]
    print(f"[parity] scal_step n={n} ncomp={nc} ppm_type={ppm}", flush=True)
    st = stepper_for(n, nc, dt, eps, ppm)
    g = st.step(to_dev(s), tuple(to_dev(x) for x in u)).cpu().numpy()
    o = pyoracle.scal_step(s, u, dt, [1.0 / n] * 3, eps, ppm_type=ppm)
    assert np.array_equal(g, o), f"{int((g != o).sum())} values differ"


@pytest.mark.parametrize("n,ppm", [(16, 1), (12, 1), (16, 2), (12, 2)])
def test_vjp_dot_product(n, ppm):
    s, u, dt, eps = setup_case(n=n, seed=3)
    n, nc = s.shape[1], s.shape[0]
    print(f"[parity] scal_step vjp dot product n={n} ppm_type={ppm}", flush=True)
    rng = np.random.default_rng(11)
    ws = rng.standard_normal(s.shape)
    wu = [rng.standard_normal(u[0].shape) for _ in range(3)]
    v = rng.standard_normal(s.shape)
    _, jw = pyoracle.scal_step(s, u, dt, [1.0 / n] * 3, eps, ws, wu, ppm_type=ppm)
    st = stepper_for(n, nc, dt, eps, ppm)
    sb, ub = st.vjp(to_dev(s), tuple(to_dev(x) for x in u), to_dev(v))
    sb = sb.cpu().numpy()
    ub = [x.cpu().numpy() for x in ub]
    lhs = float(np.sum(v * jw))
    rhs = float(np.sum(sb * ws) + sum(np.sum(ub[d] * wu[d]) for d in range(3)))
    print(f"[adjoint] <v, J w> = {lhs!r}  <J^T v, w> = {rhs!r}", flush=True)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)
    # the state part alone (w in the velocity zero)
    _, jws = pyoracle.scal_step(s, u, dt, [1.0 / n] * 3, eps, ws, None, ppm_type=ppm)
    assert abs(float(np.sum(v * jws)) - float(np.sum(sb * ws))) <= 1e-12 * abs(float(np.sum(v * jws)))


@pytest.mark.parametrize("ppm", [1, 2])
def test_conservative_step_and_dot_product(ppm):
    """The conservative form (is_conservative = 1) under a velocity with
    divergence: the GPU step equals the oracle's conservative step bit for
    bit, and the adjoint passes the dot-product test against the oracle's
    forward-mode derivative."""
    from maestroex_amd.adjoint import ScalStepper
    s, u, dt, eps = setup_case(n=16, seed=15)
    rng = np.random.default_rng(16)
    u = [x + 0.2 * rng.standard_normal(x.shape) for x in u]
    n, nc = s.shape[1], s.shape[0]
    print(f"[parity] conservative scal_step and vjp n={n} ppm_type={ppm}", flush=True)
    dx = [1.0 / n] * 3
    st = ScalStepper(n, nc, dt, eps, ppm_type=ppm, conservative=True)
    g = st.step(to_dev(s), tuple(to_dev(x) for x in u)).cpu().numpy()
    o = pyoracle.scal_step_bc(s, u, dt, dx, eps, [1, 1, 1], ppm_type=ppm, conservative=True)
    assert np.array_equal(g, o), f"{int((g != o).sum())} values differ"
    ws = rng.standard_normal(s.shape)
    wu = [rng.standard_normal(x.shape) for x in u]
    v = rng.standard_normal(s.shape)
    _, jw = pyoracle.scal_step_bc(s, u, dt, dx, eps, [1, 1, 1], ws, wu, ppm_type=ppm,
                                  conservative=True)
    sb, ub = st.vjp(to_dev(s), tuple(to_dev(x) for x in u), to_dev(v))
    sb = sb.cpu().numpy()
    ub = [x.cpu().numpy() for x in ub]
    lhs = float(np.sum(v * jw))
    rhs = float(np.sum(sb * ws) + sum(np.sum(ub[d] * wu[d]) for d in range(3)))
    print(f"[adjoint] conservative <v, J w> = {lhs!r}  <J^T v, w> = {rhs!r}", flush=True)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)


VJP_VARIANTS = [
    {"adjoint_state_fused": 0},
    {"adjoint_face_sums": 1},
    {"adjoint_pow2_index": 0},
    {"adjoint_trace_rolled": 1},
    {"adjoint_state_fused": 0, "adjoint_face_sums": 1, "adjoint_pow2_index": 0},
    {"adjoint_xcd": 1},
]
VJP_DEFAULTS = {"adjoint_state_fused": 1, "adjoint_face_sums": 0, "adjoint_pow2_index": 1,
                "adjoint_trace_rolled": 0, "adjoint_xcd": 0}


@pytest.mark.parametrize("cons", [False, True])
@pytest.mark.parametrize("variant", VJP_VARIANTS)
def test_vjp_kernel_variants_are_identical(variant, cons):
    """The adjoint's kernel arrangements (one- or three-kernel state gather,
    face adjoints stored or recomputed, mask or compare indexing) evaluate
    the same expressions: equal bit for bit, in both forms."""
    from maestroex_amd.adjoint import ScalStepper
    s, u, dt, eps = setup_case(seed=5)
    n, nc = s.shape[1], s.shape[0]
    print(f"[parity] scal_step vjp variant {variant} cons={cons}", flush=True)
    v = np.random.default_rng(12).standard_normal(s.shape)
    st = ScalStepper(n, nc, dt, eps, conservative=cons)
    args = (to_dev(s), tuple(to_dev(x) for x in u), to_dev(v))
    sb0, ub0 = st.vjp(*args)
    for key, val in variant.items():
        st.m.set_tuning(key, val)
    try:
        sb1, ub1 = st.vjp(*args)
    finally:
        for key, val in VJP_DEFAULTS.items():
            st.m.set_tuning(key, val)
    assert np.array_equal(sb0.cpu().numpy(), sb1.cpu().numpy())
    for d in range(3):
        assert np.array_equal(ub0[d].cpu().numpy(), ub1[d].cpu().numpy())


def test_autograd_function_matches_vjp():
    import torch
    from maestroex_amd.adjoint import ScalStepFn
    s, u, dt, eps = setup_case(seed=5)
    n, nc = s.shape[1], s.shape[0]
    print("[parity] scal_step autograd", flush=True)
    st = stepper_for(n, nc, dt, eps)
    sd = to_dev(s).requires_grad_(True)
    ud = [to_dev(x).requires_grad_(True) for x in u]
    tgt = to_dev(np.random.default_rng(2).standard_normal(s.shape))
    out = ScalStepFn.apply(st, sd, *ud)
    loss = 0.5 * ((out - tgt) ** 2).sum()
    loss.backward()
    g = (out - tgt).detach()
    sb, ub = st.vjp(to_dev(s), tuple(to_dev(x) for x in u), g)
    assert torch.equal(sd.grad, sb)
    for d in range(3):
        assert torch.equal(ud[d].grad, ub[d])


def rollout_loss_and_grad(st, s0, u, tgt, nsteps, every):
    from maestroex_amd.adjoint import Rollout
    sd = to_dev(s0).requires_grad_(True)
    ud = [to_dev(x).requires_grad_(True) for x in u]
    out = Rollout.apply(st, nsteps, every, sd, *ud)
    loss = 0.5 * ((out - to_dev(tgt)) ** 2).sum()
    loss.backward()
    return float(loss), sd.grad.cpu().numpy(), [x.grad.cpu().numpy() for x in ud]


def rollout_final(st, s0, u, nsteps):
    s = to_dev(s0)
    ud = tuple(to_dev(x) for x in u)
    for _ in range(nsteps):
        s = st.step(s, ud)
    return s.cpu().numpy()


def loss_difference(ap, am, tgt):
    """L(+) - L(-) of L = sum (a - t)^2 / 2, summed per cell as
    (a+ - a-)(a+ + a- - 2t)/2 so that the difference of two nearly equal
    totals is not formed."""
    return float(np.sum(0.5 * (ap - am) * (ap + am - 2.0 * tgt)))


def test_rollout_checkpointing_and_finite_differences():
    n, nc, nsteps = 16, 4, 20
    s0, u, dt, eps = setup_case(n=n, ncomp=nc, seed=7)
    tgt = np.random.default_rng(9).standard_normal(s0.shape)
    print(f"[parity] rollout n={n} ncomp={nc} steps={nsteps}", flush=True)
    st = stepper_for(n, nc, dt, eps)
    L5, gs5, gu5 = rollout_loss_and_grad(st, s0, u, tgt, nsteps, 5)
    L20, gs20, gu20 = rollout_loss_and_grad(st, s0, u, tgt, nsteps, nsteps)
    assert L5 == L20
    assert np.array_equal(gs5, gs20) and all(np.array_equal(a, b) for a, b in zip(gu5, gu20))

    h = 1e-6
    rng = np.random.default_rng(13)
    checked, kinks, worst = 0, 0, 0.0
    for which in ("s", "u"):
        for _ in range(64):
            c = rng.integers(0, nc)
            d = rng.integers(0, 3)
            k, j, i = rng.integers(0, n, size=3)

            def perturbed(delta):
                if which == "s":
                    s = s0.copy()
                    s[c, k, j, i] += delta
                    return rollout_final(st, s, u, nsteps)
                uu = [x.copy() for x in u]
                uu[d][k, j, i] += delta
                return rollout_final(st, s0, uu, nsteps)

            central = loss_difference(perturbed(h), perturbed(-h), tgt) / (2 * h)
            half = loss_difference(perturbed(h / 2), perturbed(-h / 2), tgt) / h
            g = gs5[c, k, j, i] if which == "s" else gu5[d][k, j, i]
            scale = max(abs(central), abs(g), 1e-3)
            # away from a switch the two central differences agree to O(h^2)
            if abs(central - half) > 1e-6 * scale:
                kinks += 1
                print(f"[fd] {which} point at a limiter switch: D(h) {central} D(h/2) {half}",
                      flush=True)
                continue
            err = abs(central - g) / scale
            worst = max(worst, err)
            checked += 1
            assert err <= 1e-6, f"{which} ({c},{d},{k},{j},{i}): fd {central} adjoint {g}"
    print(f"[fd] {checked} points checked, worst relative error {worst:.2e}, {kinks} at switches",
          flush=True)
    assert kinks <= 16 and checked >= 100
