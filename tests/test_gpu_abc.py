"""Learned thin absorbing boundary (cpml_abc_*, 3D FP32).

Reference for the corrector: a plain PyTorch restatement on the CPU of the
documented network (per face: the 9 scaled fields in a 4-deep box, three
3x3x3 convolutions with zero padding, torch.nn.functional.conv3d, output at
depth 0), its loss (owned outermost cells) and its gradients (torch autograd
in float64).  Tolerances:
  * FP32 training forward vs torch: rel-L2 1e-5;
  * BF16 MFMA inference vs torch: rel-L2 3e-2 (bf16 weights and activations);
  * FP32 gradients vs torch float64: rel-L2 1e-4 per face and layer (loss of
    a face: squared error over its owned outermost cells / max(target energy
    there, 1% of the box's input energy)).
A zero corrector must leave the plain step bit-identical.
"""
import numpy as np
import pytest
import torch

from seismic_cpml_amd import CpmlAbcNet, CpmlGpu, learned_abc, no_absorption

pytestmark = pytest.mark.gpu

FIELDS = ("vx", "vy", "vz", "sigmaxx", "sigmayy", "sigmazz", "sigmaxy", "sigmaxz", "sigmayz")
SCALES = learned_abc.field_scales(1.0e4)


def random_state_solver(n, seed, amp=1.0):
    """3D FP32 handle without absorption whose 9 fields are random."""
    nx, ny, nz = n
    s = CpmlGpu(3, 2, 4, nx, ny, nz, 10.0, 10.0, 10.0, 1e-4, 1, 4)
    shp = (nz, ny, nx)
    s.set_medium(np.full(shp, 2e10, np.float32), np.full(shp, 1e10, np.float32),
                 np.full(shp, 2500.0, np.float32))
    s.set_pml(no_absorption(nx), no_absorption(ny), no_absorption(nz))
    rng = np.random.default_rng(seed)
    state = {}
    for k, f in enumerate(FIELDS):
        a = (rng.standard_normal(shp) * amp * SCALES[k]).astype(np.float32)
        s.set_field(f, a)
        state[f] = a
    return s, state


def face_axes(f):
    ax = f // 2
    return ax, f % 2, (1, 2) if ax == 0 else ((0, 2) if ax == 1 else (0, 1))


def take(arr, ax, g):
    """Slice of a (nz, ny, nx) array at index g along grid axis ax -> (nb, na)."""
    if ax == 0:
        return arr[:, :, g]
    return arr[:, g, :] if ax == 1 else arr[g, :, :]


BOX = 4  # corrector input box depth; outputs at depth 0


def boxes(state, n, depth=BOX):
    out = []
    for f in range(6):
        ax, side, _ = face_axes(f)
        planes = []
        for d in range(depth):
            g = n[ax] - 1 - d if side else d
            planes.append(np.stack([take(state[name], ax, g) / SCALES[k]
                                    for k, name in enumerate(FIELDS)]))
        b = np.stack(planes, axis=1)  # (9, depth, nb, na)
        pad = np.zeros((16,) + b.shape[1:], np.float64)
        pad[:9] = b
        out.append(pad)
    return out


def torch_weight(w, f, layer):
    wl = torch.from_numpy(np.asarray(w[f, layer], np.float64)).reshape(3, 3, 3, 16, 16)
    # [kd][ka][kb][ci][co] -> [co][ci][kd][kb][ka]
    return wl.permute(4, 3, 0, 2, 1).contiguous()


def torch_forward(x, w, f):
    h = torch.from_numpy(x)[None]
    for layer in range(3):
        h = torch.nn.functional.conv3d(h, torch_weight(w, f, layer), padding=1)
    return h[0]


def owner_mask(n, f):
    ax, side, (ta, tb) = face_axes(f)
    nx, ny, nz = n
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    dist = np.stack([i, nx - 1 - i, j, ny - 1 - j, k, nz - 1 - k])
    own = np.argmin(dist, axis=0) == f
    return np.stack([take(own, ax, n[ax] - 1 - d if side else d) for d in range(4)])


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("n", [(20, 20, 20), (23, 30, 17)])
def test_corrector_forward_matches_torch(n):
    s, state = random_state_solver(n, 1)
    w = learned_abc.init_weights(2, std=0.08)
    with CpmlAbcNet(SCALES, w) as net:
        y32 = net.correction(s, use_mfma=False)
        ybf = net.correction(s, use_mfma=True)
    xs = boxes(state, n)
    for f in range(6):
        ref = torch_forward(xs[f], w, f)[:9, :1].numpy()
        assert y32[f].shape[:2] == (9, 4)
        assert rel(y32[f][:, :1], ref) < 1e-5, f
        assert rel(ybf[f][:, :1], ref) < 3e-2, f
        assert not y32[f][:, 1:].any() and not ybf[f][:, 1:].any()
    s.close()


def test_corrector_gradient_matches_torch():
    n = (18, 21, 16)
    off = (3, 2, 5)
    student, st = random_state_solver(n, 11)
    teacher, tt = random_state_solver((n[0] + 7, n[1] + 4, n[2] + 9), 12)
    w = learned_abc.init_weights(3, std=0.08)
    with CpmlAbcNet(SCALES, w) as net:
        net.train_init()
        loss, grad = net.grad(student, teacher, off)
    xs = boxes(st, n)
    tw = {k: v[off[2]:off[2] + n[2], off[1]:off[1] + n[1], off[0]:off[0] + n[0]]
          for k, v in tt.items()}
    tb = boxes(tw, n, depth=1)
    xb = boxes(st, n, depth=1)
    for f in range(6):
        W = [torch_weight(w, f, l).requires_grad_(True) for l in range(3)]
        h = torch.from_numpy(xs[f])[None]
        for l in range(3):
            h = torch.nn.functional.conv3d(h, W[l], padding=1)
        y = h[0, :9, :1]
        target = torch.from_numpy(tb[f][:9] - xb[f][:9])
        m = torch.from_numpy(owner_mask(n, f)[:1].astype(np.float64))[None]
        den = max(float(((target * m) ** 2).sum()), 1e-2 * float((xs[f] ** 2).sum()))
        lf = (((y - target) * m) ** 2).sum() / den
        lf.backward()
        assert abs(loss[f] - lf.item()) <= 1e-4 * abs(lf.item()), f
        for l in range(3):
            g_ref = W[l].grad.permute(2, 4, 3, 1, 0).reshape(27, 16, 16).numpy()
            assert rel(grad[f, l], g_ref) < 1e-4, (f, l)
    student.close()
    teacher.close()


def test_zero_corrector_leaves_plain_step_unchanged():
    m = learned_abc.layered_model(24, seed=5, f0=12.0)
    a, _, ka = learned_abc.make_run(m, 4, 0, 120)
    b, _, kb = learned_abc.make_run(m, 4, 0, 120)
    with CpmlAbcNet(SCALES) as net:
        b.set_boundary(net)
        a.run()
        b.run()
        for f in FIELDS:
            np.testing.assert_array_equal(a.field(f), b.field(f), err_msg=f)
    a.close()
    b.close()


def test_training_reduces_loss_and_learned_run_is_finite():
    torch.manual_seed(0)
    with CpmlAbcNet(learned_abc.field_scales(), learned_abc.init_weights(4, std=0.02)) as net:
        net.train_init(lr=2e-3, batch=4)
        m = learned_abc.layered_model(24, seed=7, f0=12.0, src_margin=10)
        # the Ricker pulse peaks at step 1.2 / f0 / DELTAT = 190
        teacher, _, kt = learned_abc.make_run(m, 20, 20, 420)
        student, _, ks = learned_abc.make_run(m, 4, 0, 1, source=False)
        teacher.run(1, 149)
        losses = [net.train_step(student, teacher, (16, 16, 16), it).sum()
                  for it in range(150, 421)]
        assert np.isfinite(losses).all()
        assert np.mean(losses[-40:]) < 0.5 * np.mean(losses[:40])
        teacher.close()
        student.close()
        s, _, k = learned_abc.make_run(m, 4, 0, 200)
        s.set_boundary(net)
        s.run()
        assert np.isfinite(s.field("vx")).all()
        s.set_boundary(None)
        s.close()


def test_reflection_measurement_orders_boundaries():
    """The measurement itself: a 20-cell C-PML reflects far less than the
    rigid 4-cell strip (both against the doubled-domain reference)."""
    m = learned_abc.layered_model(32, seed=9, f0=12.0)
    r, _ = learned_abc.reflection(m, 260, per_axis=5)
    assert r["cpml20"]["finite"] and r["rigid4"]["finite"]
    assert r["cpml20"]["db"] < r["rigid4"]["db"] - 10.0


def test_zero_corrector_over_thin_cpml_leaves_step_unchanged():
    """The corrector may sit on a strip with its own 4-cell C-PML; with zero
    weights the step is that thin C-PML's, bit for bit."""
    m = learned_abc.layered_model(24, seed=5, f0=12.0)
    a, _, ka = learned_abc.make_run(m, 4, 4, 120)
    b, _, kb = learned_abc.make_run(m, 4, 4, 120)
    with CpmlAbcNet(SCALES) as net:
        b.set_boundary(net)
        a.run()
        b.run()
        for f in FIELDS:
            np.testing.assert_array_equal(a.field(f), b.field(f), err_msg=f)
        assert np.abs(a.field("vx")).max() > 0
    a.close()
    b.close()


def test_training_over_thin_cpml_student_is_finite():
    """Training against a student whose strip is a 4-cell C-PML (the student
    keeps its own memory variables across samples): finite losses, and the
    trained corrector over the thin C-PML runs finite in closed loop."""
    with CpmlAbcNet(learned_abc.field_scales(), learned_abc.init_weights(4, std=0.02)) as net:
        net.train_init(lr=2e-3, batch=4)
        m = learned_abc.layered_model(24, seed=7, f0=12.0, src_margin=10)
        teacher, _, kt = learned_abc.make_run(m, 20, 20, 300)
        student, _, ks = learned_abc.make_run(m, 4, 4, 1, source=False)
        teacher.run(1, 149)
        losses = [net.train_step(student, teacher, (16, 16, 16), it).sum()
                  for it in range(150, 301)]
        assert np.isfinite(losses).all() and max(losses) > 0
        teacher.close()
        student.close()
        s, _, k = learned_abc.make_run(m, 4, 4, 200)
        s.set_boundary(net)
        s.run()
        assert np.isfinite(s.field("vx")).all()
        s.set_boundary(None)
        s.close()
