"""Parameter tables of n-state one-coordinate Morse models (the form of Coker and Xiao, J. Chem. Phys.
102, 496 (1995), NQCModels' ThreeStateMorse for n = 3) used by the n-state FSSH / Ehrenfest tests:
V_ii = D_i (1 - exp(-beta_i (x - r_i)))^2 + c_i, V_ij = A_ij exp(-a_ij (x - r_ij)^2).
Table layout: [D, beta, r, c] per state, then [A, a, r] per pair i < j in row-major order."""
import numpy as np

MASS = 20000.0


def morse_table(n, seed=0):
    """A chain of n bound Morse curves (energies within 0.3 Eh of each other for x in [2.5, 6]),
    coupled by broad Gaussians (neighbours strongest, then two apart, then every other pair), so no two
    diabats cross uncoupled and every adiabatic gap stays open."""
    rng = np.random.default_rng(seed)
    states = []
    for i in range(n):
        D = 0.02 - 0.004 * (i % 3) + 0.002 * rng.random()
        beta = 0.4 + 0.25 * (i % 2) + 0.05 * rng.random()
        r0 = 4.0 + 0.15 * i + 0.2 * rng.random()
        c = 0.02 * ((i + 1) % 2) + 0.002 * i
        states.append([D, beta, r0, c])
    pairs = []
    for i in range(n):
        for j in range(i + 1, n):
            if j == i + 1:
                pairs.append([0.004, 1.0, 3.4 + 0.3 * i + 0.1 * rng.random()])
            elif j == i + 2:
                pairs.append([0.002, 0.5, 3.6 + 0.3 * i])
            else:
                pairs.append([0.001, 0.2, 4.0 + 0.15 * (i + j)])
    return np.array(states, float).ravel().tolist() + np.array(pairs, float).ravel().tolist()


def three_state():
    """Three-state Morse test parameters (ThreeStateMorse form, model 1-like values)."""
    return [0.02, 0.4, 4.0, 0.02,
            0.02, 0.65, 4.5, 0.0,
            0.003, 0.65, 6.0, 0.02,
            0.005, 32.0, 3.40,     # (1, 2)
            0.0, 32.0, 4.8,        # (1, 3)
            0.005, 32.0, 4.97]     # (2, 3)
