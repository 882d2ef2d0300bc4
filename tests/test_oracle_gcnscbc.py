"""The checker's gcnscbc (bctype 23, src/bc.F90:7907-8137), pinned by the
characteristic relations it imposes rather than by restating it.

From a uniform interior state every LODI amplitude but the incoming wave's
is zero, so the node it writes on the face must differ from the one inside
only along the incoming characteristic (Poinsot & Lele 1992, Eqs. 33-36):
  outgoing wave l5 = 0:   d(p) + rho css d(u_n) = 0
  entropy wave l2 = 0:    d(T) = (gamma-1) T d(p) / (rho css)
  shear waves l3, l4 = 0: the tangential velocities unchanged,
with css the sound speed as the reference holds it (an INTEGER variable:
the true sound speed truncated toward zero), the pressure moved toward
pinf = 1/(gamma Ma^2) by the soft-outflow relaxation, and the density the
equation of state's thermal(p, T) (the LODI density is discarded)."""
import numpy as np
import pytest

import pyoracle as po

HM = 5
DIMS = (24, 20, 0)


def uniform_state(p, u=(0.3, 0.2, 0.05), rho=1.3, mach=0.5):
    sh = (DIMS[0] + 1 + 2 * HM, DIMS[1] + 1 + 2 * HM, 1 + 2 * HM)
    f = lambda v: np.full(sh, float(v))
    return po.prim_to_state(f(rho), f(u[0]), f(u[1]), f(u[2]), f(p), mach=mach)


def face_values(st, face):
    n0, n1 = DIMS[0], DIMS[1]
    sl = (HM + n0, slice(HM, HM + n1 + 1), HM) if face == 2 else (slice(HM, HM + n0 + 1), HM + n1, HM)
    vel = [st["vel"][..., c][sl] for c in range(3)]
    return st["rho"][sl], vel, st["prs"][sl], st["tmp"][sl]


@pytest.mark.parametrize("face", [2, 4])
def test_uniform_state_at_pinf_is_left_alone(face):
    mach = 0.5
    pinf = 1.0 / (1.4 * mach * mach)
    box = po.UwdBox(DIMS, npdc=(4, 4, 3), mach=mach)
    st = uniform_state(pinf, mach=mach)
    o = box.gcnscbc(st, face)
    r, v, p, t = face_values(o, face)
    r0, v0, p0, _ = face_values(st, face)
    assert np.allclose(p, p0, rtol=1e-13, atol=0) and np.allclose(r, r0, rtol=1e-13, atol=0)
    for c in range(3):
        assert np.allclose(v[c], v0[c], rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("face", [2, 4])
def test_only_the_incoming_wave_changes(face):
    mach = 0.5
    gamma = 1.4
    pinf = 1.0 / (gamma * mach * mach)
    rho = 1.3
    p = 1.2 * pinf
    box = po.UwdBox(DIMS, npdc=(4, 4, 3), mach=mach)
    st = uniform_state(p, rho=rho, mach=mach)
    o = box.gcnscbc(st, face)
    r, v, pn, t = face_values(o, face)
    t_in = p / rho * gamma * mach * mach
    css = int(np.sqrt(t_in) / mach)          # the reference's INTEGER sound speed
    assert css >= 1
    n = 0 if face == 2 else 1                # normal component
    tg = 1 if face == 2 else 0
    dp, dr = pn - p, r - rho
    du = v[n] - st["vel"][..., n][HM, HM, HM]
    assert np.all(np.abs(dp) > 1e-6 * p)    # the relaxation acted
    assert np.all(np.sign(dp) == -1)         # toward pinf, from above
    assert np.allclose(dp + rho * css * du, 0.0, atol=1e-9 * abs(dp).max())  # l5 = 0
    dt_ = t - t_in
    assert np.allclose(dt_, (gamma - 1.0) * t_in * dp / (rho * css), rtol=1e-9)  # l2 = 0
    assert np.all(np.abs(dr) > 0)
    for c in (tg, 2):                        # l3 = l4 = 0
        assert np.allclose(v[c], st["vel"][..., c][HM, HM, HM], rtol=1e-12, atol=1e-15)
    # the face's density is thermal(p, T) = p/T * gamma Ma^2
    assert np.allclose(r, pn / t * gamma * mach * mach, rtol=1e-14)


def test_zero_integer_sound_speed_gives_nan():
    # sqrt(T)/Ma < 1: the reference divides by zero; the checker writes NaN
    mach = 2.5
    box = po.UwdBox(DIMS, npdc=(4, 4, 3), mach=mach)
    st = uniform_state(1.0 / (1.4 * mach * mach), mach=mach)
    o = box.gcnscbc(st, 2)
    r, _, _, _ = face_values(o, 2)
    assert np.all(np.isnan(r))


def test_other_faces_untouched():
    box = po.UwdBox(DIMS, npdc=(4, 4, 3), mach=0.5)
    st = uniform_state(3.0, mach=0.5)
    for face in (1, 3, 5, 6):
        o = box.gcnscbc(st, face)
        assert np.array_equal(o["q"], st["q"])
