"""CPU restatement of the reference's eigenvector observables (test
infrastructure only: imported by tests, never by the product path).

Each function follows struct Observables in
src/observables/fermion_observables.hpp and cites the lines it restates.
Inputs: the sector's states (uint64, the reference's ascending order) and one
vector psi over them (real or complex).  Small cases only: the per-state loops
are Python.
"""
import numpy as np


def _bits(s):
    out = []
    while s:
        low = s & -s
        out.append(low.bit_length() - 1)
        s ^= low
    return out


def evaluate_cdc(k1, k2, state, highest_orbital):
    """Evaluate_CdC (src/utilities/mathematics/fermions.hpp:155-225):
    c+_{k1} c_{k2} |state> = sign |out>; returns (sign, out), sign 0 when it vanishes."""
    k1b, k2b = 1 << k1, 1 << k2
    if k2b > highest_orbital or not (state & k2b) or (not (k1b & k2b) and (state & k1b)):
        return 0, 0
    if k1b & k2b:
        return 1, state
    c = bin(state & (k2b - 1)).count("1")
    out = state ^ k2b
    c += bin(out & (k1b - 1)).count("1")
    return (-1 if c & 1 else 1), out | k1b


def density(states, psi, norb):
    """DensityFunction (fermion_observables.hpp:47-106): bin |psi_i|^2 on every
    occupied orbital of state i."""
    out = np.zeros(norb)
    for s, a in zip(states.tolist(), psi):
        p = float(np.real(a * np.conj(a)))
        for k in _bits(s):
            out[k] += p
    return out


def density_density(states, psi, norb):
    """DensityDensityFunction (:115-198): (k,k) += p for occupied k, and (k,q)
    += p for every other occupied q."""
    out = np.zeros((norb, norb))
    for s, a in zip(states.tolist(), psi):
        p = float(np.real(a * np.conj(a)))
        occ = _bits(s)
        for k in occ:
            out[k, k] += p
            for q in occ:
                if q != k:
                    out[k, q] += p
    return out


def translational_dd(states, psi, n_particles, norb, add, sub):
    """TranslationalDensityDensityFunction (:216-317): for every occupied k1 and
    every q, c+_{k1+q} c_{k1}, then for every k2 occupied in the result,
    c+_{k2-q} c_{k2}; add sign*sign2*p when the state is restored.  Divided by N."""
    nq = add.shape[1]
    out = np.zeros(nq)
    highest = 1 << (norb - 1)
    for s, a in zip(states.tolist(), psi):
        p = float(np.real(a * np.conj(a)))
        for k1 in _bits(s):
            for q in range(nq):
                s1, o1 = evaluate_cdc(int(add[k1, q]), k1, s, highest)
                if s1 == 0:
                    continue
                for k2 in _bits(o1):
                    s2, o2 = evaluate_cdc(int(sub[k2, q]), k2, o1, highest)
                    if o2 == s and s2 != 0:
                        out[q] += s1 * s2 * p
    return out / n_particles


def rotational_dd(states, psi, n_particles, inv_r60):
    """RotationalDensityDensityFunction (:334-428): for states with p > 1e-11,
    every occupied pair (k1, k2Temp) and alpha with InverseR60(k2Temp, alpha)
    == k1 adds p exp(i pi m alpha / 3) to R(m); divided by 6N."""
    out = np.zeros(6, dtype=np.complex128)
    for s, a in zip(states.tolist(), psi):
        p = float(np.real(a * np.conj(a)))
        if abs(p) <= 1e-11:
            continue
        occ = _bits(s)
        for k1 in occ:
            for k2t in occ:
                for alpha in range(6):
                    if int(inv_r60[k2t, alpha]) == k1:
                        for m in range(6):
                            out[m] += p * np.exp(1j * (np.pi / 3.0) * (m * alpha))
    return out / (6 * n_particles)


def participation_ratio(psi):
    """GetParticipationRatio (:528-563): 1 / sum |psi|^4."""
    return 1.0 / float(np.sum(np.abs(psi) ** 4))


def most_probable(states, psi, n):
    """GetMostProbable (:437-517) for one vector: the n largest |psi_i| in
    descending order.  The reference's partial quicksort leaves the order of
    equal magnitudes open; ties are taken here in ascending index order."""
    order = sorted(range(len(psi)), key=lambda i: (-abs(psi[i]), i))[:n]
    return states[order], psi[order]
