"""Closed-form pins of the oracle's contact-step logic (step_with_ipc's own
arithmetic, crates/vistio-solver/src/pd_solver.rs):

  * the CCD safety factor and backtracking, :895-932
  * the multiplier update and penalty growth, :1006-1021
  * the velocity filter (normal removal, Coulomb stick/slip, adaptive
    damping), :1036-1111
  * the global 5 % contact damping, :1115-1117 (state.rs:233-240)

Every expected value below is worked out by hand from the reference's text
(the numbers are chosen so that the arithmetic is exact or short). Each
group also plants a plausible mistake in the oracle (ref.set_mutation) and
shows that the same cases reject it, and a contact scene shows that the
oracle's step_ipc runs these very pieces (each planted mistake changes its
result).
"""
import numpy as np
import pytest

from oracle import ref


@pytest.fixture(autouse=True)
def _no_mutation():
    ref.set_mutation(0)
    yield
    ref.set_mutation(0)


def col(*v):
    return np.array(v, dtype=np.float64).reshape(3, 1)


def vf(v, n, mu_f, cd, adaptive=False, cd_max=0.5, v_thresh=0.1, inv_mass=1.0, ever=1):
    return ref.ipc_velocity_filter(v, [inv_mass], [ever], n, mu_f, cd, adaptive, cd_max, v_thresh)[:, 0]


UP = col(0.0, 1.0, 0.0)

# (velocity, normal, mu_f, cd, adaptive, cd_max, v_thresh) -> expected velocity
FILTER_CASES = [
    # slip: v_n = -4 removed; |v_t| = 3 > mu |v_n| = 2 -> v_t (1 - 2/3); then (1 - 0.1)
    ((3.0, -4.0, 0.0), UP, 0.5, 0.1, False, 0.5, 0.1, (0.9, 0.0, 0.0)),
    # stick: mu |v_n| = 4 >= |v_t| = 3 -> 0
    ((3.0, -4.0, 0.0), UP, 1.0, 0.1, False, 0.5, 0.1, (0.0, 0.0, 0.0)),
    # separating normal velocity is removed too (e = 0): v_n = +4, same friction
    ((3.0, 4.0, 0.0), UP, 0.5, 0.1, False, 0.5, 0.1, (0.9, 0.0, 0.0)),
    # no tangential velocity (|v_t| <= 1e-8): only the damping of the zero vector
    ((0.0, -4.0, 0.0), UP, 0.5, 0.1, False, 0.5, 0.1, (0.0, 0.0, 0.0)),
    # adaptive damping: after friction |v| = 1, threshold 2 -> t = 0.5, damp 0.1 + 0.5 (0.5 - 0.1) = 0.3
    ((3.0, -4.0, 0.0), UP, 0.5, 0.1, True, 0.5, 2.0, (0.7, 0.0, 0.0)),
    # adaptive, fast vertex: |v| = 1 > threshold 0.5 -> t = 0 -> base damping 0.1
    ((3.0, -4.0, 0.0), UP, 0.5, 0.1, True, 0.5, 0.5, (0.9, 0.0, 0.0)),
    # adaptive, slow vertex: |v| = 0.05, threshold 0.1 -> t = 0.5 -> damp 0.3; frictionless
    ((0.05, -1.0, 0.0), UP, 0.0, 0.1, True, 0.5, 0.1, (0.035, 0.0, 0.0)),
    # tangential along z, normal +x: v_n = 2 removed; |v_t| = 4, mu |v_n| = 1 -> (1 - 1/4); cd 0
    ((2.0, 0.0, 4.0), col(1.0, 0.0, 0.0), 0.5, 0.0, False, 0.5, 0.1, (0.0, 0.0, 3.0)),
]


@pytest.mark.parametrize("case", range(len(FILTER_CASES)))
def test_velocity_filter_closed_form(case):
    v, n, mu_f, cd, ad, cdm, vth, want = FILTER_CASES[case]
    got = vf(col(*v), n, mu_f, cd, ad, cdm, vth)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-15)


def test_velocity_filter_skips():
    """Pinned vertices, vertices never in contact and degenerate normals
    (|n|^2 < 0.5) keep their velocity; |n|^2 = 0.5 is not degenerate."""
    v = col(3.0, -4.0, 0.0)
    np.testing.assert_array_equal(vf(v, UP, 0.5, 0.1, inv_mass=0.0), v[:, 0])
    np.testing.assert_array_equal(vf(v, UP, 0.5, 0.1, ever=0), v[:, 0])
    np.testing.assert_array_equal(vf(v, col(0.5, 0.4, 0.0), 0.5, 0.1), v[:, 0])
    assert not np.array_equal(vf(v, col(0.5, 0.5, 0.0), 0.5, 0.1), v[:, 0])


def test_velocity_filter_is_rotation_equivariant():
    """Rotating velocity and normal together rotates the result: the slip
    case in a random frame gives R (0.9, 0, 0). Catches a component or index
    mix-up that the axis-aligned cases could hide."""
    rng = np.random.default_rng(3)
    for _ in range(5):
        q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        got = vf(q @ col(3.0, -4.0, 0.0), q @ UP, 0.5, 0.1)
        np.testing.assert_allclose(got, (q @ col(0.9, 0.0, 0.0))[:, 0], rtol=0, atol=1e-14)


def _filter_cases_hold():
    try:
        for v, n, mu_f, cd, ad, cdm, vth, want in FILTER_CASES:
            np.testing.assert_allclose(vf(col(*v), n, mu_f, cd, ad, cdm, vth), want, rtol=0, atol=1e-12)
    except AssertionError:
        return False
    return True


@pytest.mark.parametrize("mutation", [1, 3, 7])
def test_velocity_filter_rejects_planted_mistakes(mutation):
    """1: friction on the normal component (tangential and normal swapped);
    3: the normal velocity is kept; 7: adaptive damping blended the wrong
    way. Each fails at least one closed-form case."""
    assert _filter_cases_hold()
    ref.set_mutation(mutation)
    assert not _filter_cases_hold()


def _al_case():
    """lambda += mu grad where in contact, mu *= growth: vertex 0 in contact
    with lambda (1, 2, 3), grad (0.5, -1, 2), mu 10, growth 10 -> lambda
    (6, -8, 23), mu 100; vertex 1 not in contact keeps (4, 5, 6)."""
    lam = np.array([[1.0, 4.0], [2.0, 5.0], [3.0, 6.0]])
    grad = np.array([[0.5, 7.0], [-1.0, 7.0], [2.0, 7.0]])
    return ref.ipc_al_update([1, 0], grad, lam, 10.0, 10.0)


def test_al_update_closed_form():
    lam, mu = _al_case()
    np.testing.assert_array_equal(lam, [[6.0, 4.0], [-8.0, 5.0], [23.0, 6.0]])
    assert mu == 100.0


@pytest.mark.parametrize("mutation", [2, 6])
def test_al_update_rejects_planted_mistakes(mutation):
    """2: lambda -= mu grad; 6: mu grows before the update."""
    ref.set_mutation(mutation)
    lam, mu = _al_case()
    assert not np.array_equal(lam, [[6.0, 4.0], [-8.0, 5.0], [23.0, 6.0]])


# (ccd_alpha, grad_dot_d, mu, diff_sq) -> alpha.  Backtracking halves while
# grad_dot_d * mu > 1e-4 * diff_sq * 100 (= 0.01 here), at most 4 times, and
# stops once alpha < 0.1; a CCD bound below 1 is scaled by 0.8.
STEP_CASES = [
    ((1.0, -1.0, 1.0, 1.0), 1.0),      # barrier along the step: full step
    ((2.0, -1.0, 1.0, 1.0), 1.0),      # CCD bound above 1 clamps to 1 (no 0.8)
    ((0.5, -1.0, 1.0, 1.0), 0.4),      # CCD bound 0.5 -> 0.8 * 0.5
    ((1.0, 1.0, 1.0, 1.0), 0.0625),    # against the step: 1 -> 0.5 -> 0.25 -> 0.125 -> 0.0625 (4 halvings)
    ((0.5, 1.0, 1.0, 1.0), 0.05),      # 0.4 -> 0.2 -> 0.1 -> 0.05, then 0.05 < 0.1 stops
    ((0.15, 1.0, 1.0, 1.0), 0.06),     # 0.12 -> 0.06, stop
    ((1.0, 0.01, 1.0, 1.0), 1.0),      # exactly at the threshold: not strictly above, no halving
    ((1.0, 0.005, 4.0, 1.0), 0.0625),  # mu scales the test: 0.02 > 0.01
    ((1.0, 1.0, 1.0, 200.0), 1.0),     # a long step raises the threshold to 2
]


@pytest.mark.parametrize("case", range(len(STEP_CASES)))
def test_step_fraction_closed_form(case):
    args, want = STEP_CASES[case]
    assert ref.ipc_step_fraction(*args) == pytest.approx(want, rel=1e-15, abs=0)


@pytest.mark.parametrize("mutation", [4, 8])
def test_step_fraction_rejects_planted_mistakes(mutation):
    """4: halving when the barrier pushes along the step; 8: the 0.8 safety
    factor applied to a full step too."""
    ref.set_mutation(mutation)
    assert any(ref.ipc_step_fraction(*a) != pytest.approx(w, rel=1e-15, abs=0) for a, w in STEP_CASES)


def test_contact_damping_closed_form():
    """damp_velocities(0.05) scales every vertex (pinned ones included) by 0.95
    once any contact is active, and does nothing without contacts."""
    v = np.arange(1.0, 7.0).reshape(3, 2)
    np.testing.assert_array_equal(ref.ipc_contact_damping(v, 0), v)
    np.testing.assert_allclose(ref.ipc_contact_damping(v, 2), 0.95 * v, rtol=1e-15)
    ref.set_mutation(5)
    assert not np.allclose(ref.ipc_contact_damping(v, 2), 0.95 * v)


def _contact_scene_result(mutation):
    """A 6x6 sheet 3 cm above the ground (inside the barrier's reach, d_hat
    1e-3 m^2), sliding along x at 0.4 m/s and falling at 0.2 m/s, three IPC
    steps (ground barrier, friction, adaptive damping, backtracking; a
    negative AL tolerance runs every AL iteration, so the multipliers are
    updated): positions and velocities after the steps."""
    from oracle import mesh
    c = 6
    px, py, pz, idx = mesh.quad_grid(c, c, 0.3, 0.3)
    pz = py.copy()
    py = np.full_like(px, np.float32(0.0305))
    hg = mesh.hinges4(idx, px.size)
    pin = np.zeros(px.size, np.uint8)
    o = ref.Oracle(ref.F64, px, py, pz, idx, hg, pin, ref.params())
    v = np.zeros((3, px.size))
    v[0] = 0.4
    v[1] = -0.2
    o.set(1, v)
    ip = ref.ipc_params(ground=0.0, d_hat=1e-3, kappa=1e2, al_max_iterations=2, al_tolerance=-1.0, al_mu_growth=2.0,
                        friction_coefficient=0.4, contact_damping=0.3, adaptive_contact_damping=True)
    ref.set_mutation(mutation)
    try:
        _, st = o.step_ipc(ip, nsteps=3)
    finally:
        ref.set_mutation(0)
    return np.concatenate([o.get(0).ravel(), o.get(1).ravel()]), st


def test_step_ipc_runs_the_pinned_pieces():
    """Each planted mistake in the pieces above changes what step_ipc computes
    on a ground-contact scene that uses them all (two or more AL iterations,
    active contacts, friction, adaptive damping, a backtracked step): step_ipc calls the code the
    closed-form cases pin, not a copy of it."""
    base, st = _contact_scene_result(0)
    assert st.al_iterations >= 2 and st.active > 0
    changed = {}
    for m in range(1, 9):
        other, _ = _contact_scene_result(m)
        changed[m] = float(np.abs(other - base).max())
    print("max change of positions/velocities per planted mistake:", changed, flush=True)
    assert all(d > 0 for d in changed.values()), changed
