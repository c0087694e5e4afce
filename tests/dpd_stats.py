"""Statistics of the DPD thermostat on a channel with Poiseuille flow: the
kinetic temperature, the velocity distribution (variance and excess kurtosis
of the solvent's y and z components, which carry no mean flow), and the
density and flow profiles across the channel.  Shared by the GPU test
(tests/test_gpu_dpd_stats.py) and the script that writes the reference's
statistics (tests/golden/make_ref_dpd_stats.py)."""
import numpy as np

N_BINS = 20
EQUIL = 10000         # steps before sampling
SAMPLES = 800         # samples, every EVERY steps
EVERY = 20


def system():
    """A C1-size channel (4096 beads, box 16 x 16 x 20) with the reference's
    default-header physics: SOLVENT 3 cutoffs, DPD with the LJ cutoff as its
    range (no DPD_CUT_OFF), bending, FORCE_SWITCH_ON with a zero ramp, and a
    Poiseuille body force 0.1 along x; DPD gamma 4.5, kT 1, dt 0.005."""
    from mfa_brush_amd import configs
    from mfa_brush_amd.params import Flags
    p, t, r0, v = configs.channel_c1(thermostat=True, rc_dpd=None, bending=True, k_bend=5.0,
                                     const_force=(0.1, 0.0, 0.0))
    p = p.replace(flags=Flags(solvent=3, thermostat=True, dpd_cut_off=None, force_switch_on=True, bending=True,
                              poiseuille=True), r_2_min_init=0.0, r_2_min_time=0.0)
    return p, t, r0, v


class Stats:
    def __init__(self, p, a_type):
        self.p, self.t = p, np.asarray(a_type)
        self.free = self.t != 1                      # heads are frozen
        self.sol = self.t == 3
        self.m = np.asarray(p.mass_type)[self.t - 1]
        self.n = 0
        self.kt = []
        self.mom = np.zeros(5)                      # count, sum v, v^2, v^3, v^4 of the solvent's y and z components
        self.rho = np.zeros((2, N_BINS))
        self.vx = np.zeros(N_BINS)
        self.cnt = np.zeros(N_BINS)

    def add(self, r0, v):
        p = self.p
        self.n += 1
        self.kt.append(float((self.m[self.free, None] * v[self.free] ** 2).sum() / (3 * self.free.sum())))
        w = v[self.sol][:, 1:].ravel()
        self.mom += [w.size, w.sum(), (w ** 2).sum(), (w ** 3).sum(), (w ** 4).sum()]
        zb = np.clip((r0[:, 2] / p.boundary[2] * N_BINS).astype(int), 0, N_BINS - 1)
        self.rho[0] += np.bincount(zb[~self.sol], minlength=N_BINS)
        self.rho[1] += np.bincount(zb[self.sol], minlength=N_BINS)
        self.vx += np.bincount(zb[self.sol], weights=v[self.sol, 0], minlength=N_BINS)
        self.cnt += np.bincount(zb[self.sol], minlength=N_BINS)

    def result(self):
        p = self.p
        c, s1, s2, s3, s4 = self.mom
        mu = s1 / c
        var = s2 / c - mu ** 2
        m4 = s4 / c - 4 * mu * s3 / c + 6 * mu ** 2 * s2 / c - 3 * mu ** 4
        dz = p.boundary[2] / N_BINS
        area = p.boundary[0] * p.boundary[1]
        return dict(kt=float(np.mean(self.kt)), var_yz=float(var), kurt_yz=float(m4 / var ** 2 - 3.0),
                    rho=self.rho / (self.n * area * dz), vx=self.vx / np.maximum(self.cnt, 1),
                    sol_count=self.cnt / self.n)


def run(stepper, p, a_type, r0, v):
    """Two halves of the sampling window (Stats of each) from a stepper with
    md_steps(n) and state() -> (r0, v) in the reference's bead order."""
    stepper.md_steps(EQUIL)
    halves = []
    for _ in range(2):
        st = Stats(p, a_type)
        for _ in range(SAMPLES // 2):
            stepper.md_steps(EVERY)
            st.add(*stepper.state())
        halves.append(st.result())
    return halves


def merge(a, b):
    return {k: 0.5 * (np.asarray(a[k]) + np.asarray(b[k])) for k in a}


def l1(x, y):
    return float(np.abs(np.asarray(x) - np.asarray(y)).sum() / np.abs(np.asarray(y)).sum())
