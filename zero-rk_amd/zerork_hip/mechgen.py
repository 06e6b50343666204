"""Seeded synthetic CHEMKIN mechanism and reactor-state generators.

The workloads of this project are synthetic chemistry: element-balanced random
reaction sets written as ordinary CHEMKIN-II mechanism and 80-column NASA-7
thermo files, so that every parser (this package's, the oracle's and the
reference's CKReader, src/CKconverter/CKParser.cpp:861-1046 for thermo cards,
:1094-1110 for units) reads the same text.  Every generator takes an explicit
seed and is deterministic for it.

Mechanism recipe (all reactions reversible, reverse rate from Keq):
  * elementary bimolecular / third-body (+M) / Troe falloff (+M) mix, each
    reaction written in its exothermic direction (the random activation
    energy applies downhill; the reverse gets Ea + dH through Keq),
  * 1-2 reactants and 1-2 products (up to ``max_participants``), element-balanced,
  * log10 A ~ U[8,14] (cgs), beta ~ U[-1,2], Ea ~ U[0,40] kcal/mol,
  * third-body efficiencies ~ U[0.5,3],
  * NASA-7 thermo perturbed from a constant cp ~ U[3.5,6] R (atoms ~2.5 R),
    continuous in cp, h and s at Tmid = 1000 K.
"""
from __future__ import annotations

import itertools
import math
import os
from dataclasses import dataclass, field

import numpy as np

RU_CAL = 1.98720425864083  # cal/mol/K (= 8.314462618 / 4.184)

# Known species: formula -> (H_f298 [kcal/mol], S298 [cal/mol/K]); values are
# textbook-order thermochemistry used only to make equilibrium constants sane.
_KNOWN = {
    "H": (52.1, 27.4), "H2": (0.0, 31.2), "O": (59.6, 38.5), "O2": (0.0, 49.0),
    "OH": (8.9, 43.9), "H2O": (-57.8, 45.1), "HO2": (3.0, 54.8),
    "H2O2": (-32.5, 55.7), "O3": (34.1, 57.1), "N2": (0.0, 45.8),
    "AR": (0.0, 37.0), "N": (113.0, 36.6), "NO": (21.6, 50.4),
    "NO2": (8.1, 57.4), "N2O": (19.6, 52.6), "NH": (85.2, 43.3),
    "NH2": (45.1, 46.5), "NH3": (-11.0, 46.0), "HNO": (25.6, 52.8),
    "NNH": (59.6, 53.6), "HCN": (32.3, 48.2), "CN": (104.0, 48.4),
    "NCO": (30.5, 55.5), "HNCO": (-28.0, 57.0), "C": (171.3, 37.8),
    "CH": (142.5, 43.7), "CH2": (93.0, 46.3), "CH3": (35.0, 46.4),
    "CH4": (-17.9, 44.5), "CO": (-26.4, 47.2), "CO2": (-94.1, 51.1),
    "HCO": (10.4, 53.7), "CH2O": (-26.0, 52.3), "CH3O": (4.1, 54.6),
    "CH2OH": (-3.9, 58.4), "CH3OH": (-48.1, 57.3), "C2H": (135.0, 49.6),
    "C2H2": (54.2, 48.0), "C2H3": (71.0, 55.9), "C2H4": (12.5, 52.4),
    "C2H5": (28.9, 60.1), "C2H6": (-20.0, 54.9), "HCCO": (42.4, 60.7),
    "CH2CO": (-11.4, 57.8), "HCCOH": (20.0, 58.0), "CH3CHO": (-39.7, 63.2),
    "CH2CHO": (3.0, 64.0), "C2H5OH": (-56.2, 67.5), "C3H8": (-25.0, 64.5),
    "C3H6": (4.9, 63.8), "C3H4": (45.5, 58.3), "C3H3": (81.0, 61.5),
    "C3H5": (40.0, 63.0), "NC3H7": (24.0, 69.0), "IC3H7": (21.5, 69.2),
}

ATOM_H = {"H": 52.1, "O": 59.6, "N": 113.0, "C": 171.3, "AR": 0.0}


def parse_formula(name: str) -> dict:
    """Element counts of a formula-like species name (C2H5OH, IC3H7, AR)."""
    s = name.upper()
    for pre in ("IC", "NC", "SC", "TC"):
        if s.startswith(pre) and len(s) > 2 and s[2].isdigit():
            s = s[1:]
            break
    out: dict = {}
    i = 0
    while i < len(s):
        if s.startswith("AR", i):
            el, i = "AR", i + 2
        else:
            el, i = s[i], i + 1
        j = i
        while j < len(s) and s[j].isdigit():
            j += 1
        n = int(s[i:j]) if j > i else 1
        i = j
        out[el] = out.get(el, 0) + n
    return out


@dataclass
class Reaction:
    reac: list
    prod: list
    kind: str  # "elem" | "tb" | "falloff"
    A: float
    beta: float
    Ea: float  # cal/mol
    eff: dict = field(default_factory=dict)
    low: tuple = ()
    troe: tuple = ()


@dataclass
class Mechanism:
    elements: list
    species: list
    comp: list           # list of element-count dicts
    thermo: list         # list of (low[7], high[7]) coefficient tuples
    reactions: list
    seed: int
    majors: list

    def mech_text(self) -> str:
        L = [f"! synthetic mechanism, seed {self.seed}, {len(self.species)} species,"
             f" {len(self.reactions)} reactions", "ELEMENTS"]
        L.append(" ".join(self.elements))
        L.append("END")
        L.append("SPECIES")
        for i in range(0, len(self.species), 8):
            L.append(" ".join(self.species[i:i + 8]))
        L.append("END")
        L.append("REACTIONS CAL/MOLE")
        for r in self.reactions:
            lhs = " + ".join(r.reac)
            rhs = " + ".join(r.prod)
            if r.kind == "tb":
                lhs += " + M"
                rhs += " + M"
            elif r.kind == "falloff":
                lhs += " (+M)"
                rhs += " (+M)"
            L.append(f"{lhs} <=> {rhs}   {r.A:.10E} {r.beta:.6f} {r.Ea:.4f}")
            if r.kind == "falloff":
                L.append(f"   LOW / {r.low[0]:.10E} {r.low[1]:.6f} {r.low[2]:.4f} /")
                L.append("   TROE / " + " ".join(f"{v:.6E}" for v in r.troe) + " /")
            if r.eff:
                L.append("   " + " ".join(f"{k}/{v:.4f}/" for k, v in r.eff.items()))
        L.append("END")
        return "\n".join(L) + "\n"

    def therm_text(self) -> str:
        L = ["THERMO ALL", "   300.000  1000.000  5000.000"]
        for name, comp, (lo, hi) in zip(self.species, self.comp, self.thermo):
            els = "".join(f"{e:<2s}{n:>3d}" for e, n in comp.items())
            els = els.ljust(20)
            l1 = f"{name:<16s}{'SYNTH':<8s}{els}G{300.0:10.3f}{5000.0:10.3f}{1000.0:8.2f}"
            L.append(l1.ljust(79) + "1")
            c = list(hi) + list(lo)
            fmt = lambda v: f"{v:15.8E}"
            L.append("".join(fmt(v) for v in c[0:5]).ljust(79) + "2")
            L.append("".join(fmt(v) for v in c[5:10]).ljust(79) + "3")
            L.append("".join(fmt(v) for v in c[10:14]).ljust(79) + "4")
        L.append("END")
        return "\n".join(L) + "\n"

    def write(self, directory: str, stem: str) -> tuple:
        os.makedirs(directory, exist_ok=True)
        mp = os.path.join(directory, stem + ".mech")
        tp = os.path.join(directory, stem + ".therm")
        with open(mp, "w") as f:
            f.write(self.mech_text())
        with open(tp, "w") as f:
            f.write(self.therm_text())
        return mp, tp


def _nasa_coeffs(rng, cp0: float, hf_kcal: float, s298: float, atom: bool):
    """Low/high NASA-7 sets, continuous in cp, h, s at 1000 K."""
    Tm, Tr = 1000.0, 298.15
    if atom:
        lo = [cp0, 0.0, 0.0, 0.0, 0.0]
        hi = [cp0, 0.0, 0.0, 0.0, 0.0]
    else:
        lo = [cp0, rng.uniform(0.0, 1.5e-3), rng.uniform(-4e-7, 0.0), 0.0, 0.0]
        hi = [0.0, rng.uniform(0.0, 6e-4), rng.uniform(-1e-7, 0.0), 0.0, 0.0]
    cp = lambda a, T: a[0] + a[1] * T + a[2] * T**2 + a[3] * T**3 + a[4] * T**4
    hint = lambda a, T: a[0] * T + a[1] * T**2 / 2 + a[2] * T**3 / 3 + a[3] * T**4 / 4 + a[4] * T**5 / 5
    sint = lambda a, T: a[0] * math.log(T) + a[1] * T + a[2] * T**2 / 2 + a[3] * T**3 / 3 + a[4] * T**4 / 4
    if not atom:
        hi[0] = cp(lo, Tm) - (hi[1] * Tm + hi[2] * Tm**2)
    a6lo = hf_kcal * 1000.0 / RU_CAL - hint(lo, Tr)
    a7lo = s298 / RU_CAL - sint(lo, Tr)
    a6hi = hint(lo, Tm) + a6lo - hint(hi, Tm)
    a7hi = sint(lo, Tm) + a7lo - sint(hi, Tm)
    return tuple(lo + [a6lo, a7lo]), tuple(hi + [a6hi, a7hi])


def _h298(low):
    """h(298.15 K) in J/mol-equivalent units (R * T * h/RT) of a low-range set."""
    T = 298.15
    a = low
    h_rt = a[0] + a[1] * T / 2 + a[2] * T**2 / 3 + a[3] * T**3 / 4 + a[4] * T**4 / 5 + a[5] / T
    return h_rt * T


def _species_pool(elements):
    names = [n for n in _KNOWN if set(parse_formula(n)) <= set(elements)]
    return names


def _extra_species(rng, elements, n, taken, cmax=12):
    """Formula species beyond the known table (CxHyOz(Nw)) with additive thermo."""
    out = []
    hasC, hasN = "C" in elements, "N" in elements
    tries = 0
    while len(out) < n and tries < 200000:
        tries += 1
        c = int(rng.integers(1, cmax + 1)) if hasC else 0
        h = int(rng.integers(0, 2 * c + 3)) if hasC else int(rng.integers(1, 5))
        o = int(rng.integers(0, 4))
        w = int(rng.integers(0, 2)) if hasN and rng.random() < 0.15 else 0
        if c + h + o + w < 2:
            continue
        name = (f"C{c}" if c > 1 else ("C" if c == 1 else "")) + \
               (f"H{h}" if h > 1 else ("H" if h == 1 else "")) + \
               (f"O{o}" if o > 1 else ("O" if o == 1 else "")) + \
               (f"N{w}" if w > 1 else ("N" if w == 1 else ""))
        if name in taken or name in out:
            continue
        out.append(name)
    return out


def _thermo_for(rng, name, comp):
    natoms = sum(comp.values())
    if name in _KNOWN:
        hf, s = _KNOWN[name]
        hf += rng.uniform(-2.0, 2.0)
    else:
        hf = sum(ATOM_H[e] * k for e, k in comp.items()) - 95.0 * (natoms - 1) + rng.uniform(-10, 10)
        s = 25.0 + 9.0 * natoms + rng.uniform(-3, 3)
    atom = natoms == 1
    cp0 = rng.uniform(2.45, 2.55) if atom else rng.uniform(3.5, 6.0)
    return _nasa_coeffs(rng, cp0, hf, s, atom)


def _enumerate_pairs(rng, evec, reactive, n_el, n_tb, n_fo, major_idx, max_participants):
    """Element-balanced reaction sides by enumeration of every species set of up
    to max_participants participants (small mechanisms)."""
    groups: dict = {}
    for size in range(1, max_participants + 1):
        for ms in itertools.combinations_with_replacement(reactive, size):
            key = tuple(evec[list(ms)].sum(axis=0))
            groups.setdefault(key, []).append(ms)
    # candidate pairs by reaction class
    bimol, assoc = [], []  # (reac, prod)
    for key, lst in groups.items():
        if len(lst) < 2:
            continue
        for a, b in itertools.combinations(lst, 2):
            if set(a) & set(b):
                continue
            if len(a) >= 2 and len(b) >= 2:
                bimol.append((a, b))
            elif (len(a) == 1) != (len(b) == 1):
                assoc.append((a, b) if len(a) > len(b) else (b, a))
    if len(bimol) < n_el or len(assoc) < n_fo + n_tb:
        raise ValueError(f"species set too small: {len(bimol)} bimolecular, {len(assoc)} "
                         f"association candidates for {n_el} + {n_tb + n_fo} reactions")
    # grow a connected network: a candidate is accepted only when one of its
    # sides consists of species already reachable from the initial majors, so
    # every reaction can fire from a mixture the reactors actually contain
    reach = set(major_idx)
    want = {"elem": n_el, "tb": n_tb, "falloff": n_fo}
    pools = {"elem": bimol, "tb": assoc, "falloff": assoc}
    perms = {"elem": list(rng.permutation(len(bimol))), "assoc": list(rng.permutation(len(assoc)))}
    used_assoc = set()
    kinds, pairs = [], []
    schedule = list(rng.permutation(["elem"] * n_el + ["tb"] * n_tb + ["falloff"] * n_fo))
    for kind in schedule:
        pool = pools[kind]
        key = "elem" if kind == "elem" else "assoc"
        perm = perms[key]
        pick = None
        for pos, ci in enumerate(perm):
            if key == "assoc" and ci in used_assoc:
                continue
            a, b = pool[ci]
            if set(a) <= reach or set(b) <= reach:
                pick = pos
                break
        if pick is None:  # nothing reachable left: take the next candidate
            pick = next(p for p, ci in enumerate(perm) if key != "assoc" or ci not in used_assoc)
        ci = perm.pop(pick)
        if key == "assoc":
            used_assoc.add(ci)
        a, b = pool[ci]
        reach.update(a)
        reach.update(b)
        kinds.append(kind)
        pairs.append((a, b))
    want.clear()
    return pairs, kinds


def _sample_pairs(rng, evec, reactive, n_el, n_tb, n_fo, major_idx):
    """Element-balanced reaction sides for large species sets (thousands of
    species), sampled instead of enumerated: one side is drawn from species
    already reachable from the initial majors (so every reaction can fire from
    a mixture the reactors contain), the other from the species or species
    pairs with the same element totals; no reaction appears twice."""
    R = np.asarray(reactive, dtype=np.int64)
    base = int(evec.max()) * 2 + 1
    key_all = evec @ (base ** np.arange(evec.shape[1], dtype=np.int64))
    ii, jj = np.triu_indices(len(R))
    pa, pb = R[ii], R[jj]
    k2 = key_all[pa] + key_all[pb]
    order = np.argsort(k2, kind="stable")
    uniq, start, count = np.unique(k2[order], return_index=True, return_counts=True)
    single: dict = {}
    for s in R:
        single.setdefault(int(key_all[s]), []).append(int(s))

    def pair_group(k):
        pos = int(np.searchsorted(uniq, k))
        if pos < len(uniq) and uniq[pos] == k:
            return order[start[pos]:start[pos] + count[pos]]
        return None

    rset = set(int(x) for x in R)
    reach = [m for m in major_idx if m in rset]
    in_reach = set(reach)
    seen = set()
    kinds, pairs = [], []
    schedule = list(rng.permutation(["elem"] * n_el + ["tb"] * n_tb + ["falloff"] * n_fo))
    for kind in schedule:
        cand = None
        for _ in range(2000):
            if kind == "elem":
                x, y = sorted(int(v) for v in rng.choice(reach, 2))
                g = pair_group(int(key_all[x] + key_all[y]))
                if g is None:
                    continue
                q = int(g[rng.integers(len(g))])
                c, d = int(pa[q]), int(pb[q])
                if {c, d} & {x, y}:
                    continue
                cand = ((x, y), (c, d))
            elif rng.random() < 0.5:  # a reachable species and a pair of its elements
                s = int(rng.choice(reach))
                g = pair_group(int(key_all[s]))
                if g is None:
                    continue
                q = int(g[rng.integers(len(g))])
                c, d = int(pa[q]), int(pb[q])
                if s in (c, d):
                    continue
                cand = ((c, d), (s,))
            else:  # a reachable pair and the species of its elements
                x, y = sorted(int(v) for v in rng.choice(reach, 2))
                lst = single.get(int(key_all[x] + key_all[y]))
                if not lst:
                    continue
                s = lst[int(rng.integers(len(lst)))]
                if s in (x, y):
                    continue
                cand = ((x, y), (s,))
            sig = frozenset((cand[0], cand[1]))
            if sig in seen:
                cand = None
                continue
            seen.add(sig)
            break
        if cand is None:
            raise ValueError(f"no element-balanced candidate left for a {kind} reaction")
        for v in cand[0] + cand[1]:
            if v not in in_reach:
                in_reach.add(v)
                reach.append(v)
        kinds.append(kind)
        pairs.append(cand)
    return pairs, kinds


def generate(n_species: int, n_reactions: int, seed: int, elements=("H", "O", "N"),
             frac_third_body: float = 0.25, frac_falloff: float = 0.15,
             inert=("N2", "AR"), majors=None, max_participants: int = 2,
             base=None) -> Mechanism:
    """Seeded synthetic element-balanced CHEMKIN mechanism."""
    rng = np.random.default_rng(seed)
    elements = [e.upper() for e in elements]
    pool = _species_pool(elements)
    if base is not None:
        chosen = [s for s in base]
    else:
        chosen = []
    for s in pool:
        if len(chosen) >= n_species:
            break
        if s not in chosen:
            chosen.append(s)
    large = n_species > 400
    if len(chosen) < n_species:
        chosen += _extra_species(rng, elements, n_species - len(chosen), set(chosen),
                                 cmax=24 if large else 12)
    species = chosen[:n_species]
    comp = [parse_formula(s) for s in species]
    # element order as in the ELEMENTS list
    comp = [{e: c[e] for e in elements if e in c} for c in comp]
    thermo = [_thermo_for(rng, s, c) for s, c in zip(species, comp)]

    reactive = [i for i, s in enumerate(species) if s not in inert]
    evec = np.array([[c.get(e, 0) for e in elements] for c in comp], dtype=np.int64)
    n_fo = int(round(frac_falloff * n_reactions))
    n_tb = int(round(frac_third_body * n_reactions))
    n_el = n_reactions - n_fo - n_tb
    major_idx = [species.index(x) for x in (majors or species[:4]) if x in species]
    if large:
        pairs, kinds = _sample_pairs(rng, evec, reactive, n_el, n_tb, n_fo, major_idx)
    else:
        pairs, kinds = _enumerate_pairs(rng, evec, reactive, n_el, n_tb, n_fo, major_idx,
                                       max_participants)
    order = rng.permutation(n_reactions)
    # standard enthalpy of formation at 298.15 K from the generated NASA-7 data
    h298 = np.array([_h298(lo) for lo, _ in thermo])
    reactions = [_make_reaction(rng, pairs[idx], kinds[idx], species, h298, species)
                 for idx in order]
    if majors is None:
        majors = species[:4]
    return Mechanism(elements, species, comp, thermo, reactions, seed, list(majors))


def _make_reaction(rng, pair, kind, species, h298, eff_species, bounded=False):
    """One reaction with the recipe's random rate parameters.  It is written in
    its exothermic direction: the random activation energy then belongs to the
    downhill direction and the reverse (from Keq) carries Ea + dH, as in
    physical mechanisms.  Third-body efficiencies go to 1-4 species drawn from
    eff_species.  bounded: the 10^U[8,14] range applies to A (1000 K)^beta,
    the temperature factor at combustion temperatures, instead of A alone
    (with beta up to 2 an unbounded A reaches 1e20 at 1000 K, far above the
    gas-kinetic limit)."""
    a, b = pair
    if h298[list(b)].sum() > h298[list(a)].sum():
        a, b = b, a
    A = 10.0 ** rng.uniform(8.0, 14.0)
    beta = rng.uniform(-1.0, 2.0)
    if bounded:
        A /= 1000.0 ** beta
    Ea = rng.uniform(0.0, 40.0) * 1000.0
    r = Reaction([species[i] for i in a], [species[i] for i in b], kind, A, beta, Ea)
    if kind in ("tb", "falloff"):
        ne = int(rng.integers(1, 5))
        for j in rng.choice(len(eff_species), size=min(ne, len(eff_species)), replace=False):
            r.eff[eff_species[j]] = float(rng.uniform(0.5, 3.0))
    if kind == "falloff":
        r.low = (A * 10.0 ** rng.uniform(3.0, 7.0), rng.uniform(-2.0, 1.0),
                 rng.uniform(0.0, 40.0) * 1000.0)
        alpha = rng.uniform(0.1, 0.9)
        t3 = rng.uniform(50.0, 500.0)
        t1 = rng.uniform(500.0, 5000.0)
        if rng.random() < 0.5:
            r.troe = (alpha, t3, t1, rng.uniform(1000.0, 10000.0))
        else:
            r.troe = (alpha, t3, t1)
    return r


# ----------------------------------------------------------------------------
# Large mechanisms: a core of small species and families of large ones
# ----------------------------------------------------------------------------
# The pool of small species through which large species react (abstractors,
# leaving groups and fragments), as in the fuel-specific parts of real large
# mechanisms (alkyl + O2, H abstraction by H/O/OH/HO2/CH3, beta scission).
_POOL = ["H", "H2", "O", "O2", "OH", "H2O", "HO2", "H2O2", "CH3", "CH4", "CO", "CO2",
         "CH2O", "HCO", "CH3O", "C2H5", "C2H4", "C2H6", "NC3H7", "C3H6"]


def _family_members(rng, n_c, size, fam):
    """Compositions C{n}H{h}O{o} of one family of large species (isomers of one
    carbon skeleton): the alkane, alkyl radicals and the olefin first, then
    random members of the lattice 2n-3 <= h <= 2n+2, 0 <= o <= 4, with up to
    three isomers of a composition."""
    comps = [(n_c, 2 * n_c + 2, 0), (n_c, 2 * n_c + 1, 0), (n_c, 2 * n_c + 1, 0),
             (n_c, 2 * n_c, 0)]
    lattice = [(n_c, h, o) for h in range(2 * n_c - 3, 2 * n_c + 3) for o in range(0, 5)]
    while len(comps) < size:
        c = lattice[int(rng.integers(len(lattice)))]
        if comps.count(c) < 3:
            comps.append(c)
    out, seen = [], {}
    for c in comps:
        k = seen.get(c, 0)
        seen[c] = k + 1
        formula = f"C{c[0]}H{c[1]}" + (f"O{c[2]}" if c[2] > 1 else ("O" if c[2] == 1 else ""))
        out.append((f"{formula}-{fam:03d}{chr(65 + k)}", {"C": c[0], "H": c[1], "O": c[2]}))
    return out


def generate_families(n_species: int, n_reactions: int, seed: int, n_core: int = 120,
                      core_reactions: int = 600, frac_third_body: float = 0.25,
                      frac_falloff: float = 0.10, n_fuels: int = 4,
                      family_size=(24, 40), carbon=(5, 16),
                      class_weights=(0.45, 0.25, 0.10, 0.20)) -> Mechanism:
    """Seeded synthetic large mechanism with the structure of real ones: a core
    of small species (the known table, then random formulas of at most 3 C)
    whose reactions are sampled like the small mechanisms', and families of
    large C/H/O species, each the isomer lattice of one carbon skeleton.  A
    family member reacts only with members of its own family (H transfer
    A + X <=> B + Y and decomposition A <=> B + Y with X, Y from the radical
    pool, isomerisation A <=> B); a family's two alkyl radicals also split into
    a pool fragment and the alkane, an alkyl or the olefin of the first family
    of a smaller carbon count (beta scission down a ladder of skeletons).
    Elimination therefore fills families and the core, not the whole matrix.
    Kinds and thermo follow the module recipe, rate parameters too except
    that A (1000 K)^beta, not A, is drawn from 10^U[8,14] (_make_reaction,
    bounded); third-body efficiencies go to core species.  The majors are the core's eight base species and the alkanes of
    the n_fuels largest families."""
    rng = np.random.default_rng(seed)
    elements = ["C", "H", "O", "N", "AR"]
    core = list(_C2_BASE)
    for s in _species_pool(elements):
        if len(core) >= n_core:
            break
        if s not in core:
            core.append(s)
    core += _extra_species(rng, elements, n_core - len(core), set(core), cmax=3)
    # families until the species count is reached
    fams, n_left, fam = [], n_species - len(core), 0
    while n_left > 0:
        size = int(rng.integers(family_size[0], family_size[1] + 1))
        if n_left - size < family_size[0]:
            size = n_left
        n_c = int(rng.integers(carbon[0], carbon[1] + 1))
        fams.append((n_c, _family_members(rng, n_c, size, fam)))
        n_left -= size
        fam += 1
    species = list(core)
    comp = [parse_formula(x) for x in core]
    comp = [{e: c[e] for e in elements if e in c} for c in comp]
    members = []
    for n_c, mem in fams:
        idx = []
        for name, c in mem:
            idx.append(len(species))
            species.append(name)
            comp.append(c)
        members.append(idx)
    thermo = [_thermo_for(rng, x, c) for x, c in zip(species, comp)]
    h298 = np.array([_h298(lo) for lo, _ in thermo])
    evec = np.array([[c.get(e, 0) for e in elements] for c in comp], dtype=np.int64)
    key = [tuple(v) for v in evec]
    pool = [species.index(x) for x in _POOL]
    single, diff = {}, {}
    for y in pool:
        single.setdefault(key[y], []).append(y)
        for x in pool:
            if x != y:
                d = tuple(evec[y] - evec[x])
                diff.setdefault(d, []).append((x, y))
    # the first family of each carbon count is the ladder that larger
    # families' alkyl radicals scission into (alkane, alkyls, olefin)
    ladder = {}
    for f, (n_c, _) in enumerate(fams):
        ladder.setdefault(n_c, f)
    by_key = {}
    for f, idx in enumerate(members):
        for i in idx:
            by_key.setdefault(key[i], []).append((f, i))
    # core reactions: element-balanced sides sampled over the core
    reactive = [i for i, x in enumerate(core) if x not in ("N2", "AR")]
    n_fo_c = int(round(frac_falloff * core_reactions))
    n_tb_c = int(round(frac_third_body * core_reactions))
    major_core = [species.index(x) for x in _C2_BASE]
    pairs, kinds = _sample_pairs(rng, evec[:len(core)], reactive,
                                 core_reactions - n_fo_c - n_tb_c, n_tb_c, n_fo_c, major_core)
    seen = {frozenset((a, b)) for a, b in pairs}
    # family reactions: candidates by class, sampled per family in proportion
    # to its size; every member takes part in at least one reaction
    n_fam_rx = n_reactions - core_reactions
    n_fam_sp = sum(len(m) for m in members)
    n_fo = int(round(frac_falloff * n_reactions)) - n_fo_c
    n_tb = int(round(frac_third_body * n_reactions)) - n_tb_c
    fam_pairs, fam_dis = [], []  # (reac, prod), dissociation/association class?
    for f, idx in enumerate(members):
        n_c = fams[f][0]
        target = int(round(n_fam_rx * len(idx) / n_fam_sp)) if f < len(members) - 1 else \
            n_fam_rx - len(fam_pairs)
        cands = {"transfer": [], "dissoc": [], "isomer": [], "scission": []}
        for a in idx:
            for b in idx:
                if a == b:
                    continue
                d = tuple(evec[a] - evec[b])
                if not any(d):
                    if a < b:
                        cands["isomer"].append(((a,), (b,)))
                    continue
                for y in single.get(d, ()):
                    cands["dissoc"].append(((a,), (b, y)))
                if a < b:
                    for x, y in diff.get(d, ()):
                        cands["transfer"].append(((a, x), (b, y)))
            if a not in idx[1:3]:
                continue
            for y in pool:
                k = tuple(evec[a] - evec[y])
                for g, b in by_key.get(k, ()):
                    if ladder.get(fams[g][0]) == g and fams[g][0] < n_c and \
                            b in members[g][:4]:
                        cands["scission"].append(((a,), (b, y)))
        weights = dict(zip(("transfer", "dissoc", "isomer", "scission"), class_weights))
        order = {c: list(rng.permutation(len(v))) for c, v in cands.items()}
        chosen = []

        def take(cls, must=None):
            lst, perm = cands[cls], order[cls]
            for pos, ci in enumerate(perm):
                r = lst[ci]
                if must is not None and must not in r[0] + r[1]:
                    continue
                sig = frozenset(r)
                if sig in seen:
                    continue
                seen.add(sig)
                perm.pop(pos)
                chosen.append((r, cls in ("dissoc", "scission")))
                return True
            return False

        for a in rng.permutation(idx):
            if any(a in r[0] + r[1] for r, _ in chosen):
                continue
            for cls in rng.permutation(list(weights)):
                if take(str(cls), int(a)):
                    break
        classes = list(weights)
        p = np.array([weights[c] for c in classes])
        while len(chosen) < target:
            live = [i for i, c in enumerate(classes) if order[c]]
            if not live:
                raise ValueError(f"family {f}: candidates exhausted")
            q = p[live] / p[live].sum()
            take(classes[live[int(rng.choice(len(live), p=q))]])
        for r, dis in chosen:
            fam_pairs.append(r)
            fam_dis.append(dis)
    # kinds: third-body and fall-off among the association/dissociation class
    dis_idx = [i for i, d in enumerate(fam_dis) if d]
    if len(dis_idx) < n_tb + n_fo:
        raise ValueError("too few dissociation reactions for the third-body/fall-off mix")
    pick = rng.permutation(dis_idx)
    fam_kinds = ["elem"] * len(fam_pairs)
    for i in pick[:n_tb]:
        fam_kinds[i] = "tb"
    for i in pick[n_tb:n_tb + n_fo]:
        fam_kinds[i] = "falloff"
    pairs = list(pairs) + fam_pairs
    kinds = list(kinds) + fam_kinds
    for (a, b) in pairs:
        assert np.array_equal(evec[list(a)].sum(axis=0), evec[list(b)].sum(axis=0))
    order_rx = rng.permutation(len(pairs))
    eff_species = list(core)
    reactions = [_make_reaction(rng, pairs[i], kinds[i], species, h298, eff_species,
                                bounded=True) for i in order_rx]
    fuels = [species[members[f][0]] for f in
             sorted(range(len(fams)), key=lambda f: (-fams[f][0], f))[:n_fuels]]
    return Mechanism(elements, species, comp, thermo, reactions, seed,
                     list(_C2_BASE) + fuels)


# ----------------------------------------------------------------------------
# Workload configurations (mechanism + reactor state distributions)
# ----------------------------------------------------------------------------
_C2_BASE = ["H2", "O2", "N2", "H2O", "CO", "CO2", "CH4", "AR"]


WORKLOADS = ("10sp", "53sp", "160sp", "2000sp")


def workload_mechanism(name: str, seed: int | None = None) -> Mechanism:
    """Seeded synthetic mechanism of a named workload: "10sp" (H/O/N, 27 reactions),
    "53sp" (C/H/O/N/Ar, 325 reactions), "160sp" (C/H/O/N/Ar, 900 reactions) or
    "2000sp" (C/H/O/N/Ar, 8,000 reactions: a 120-species core and families of
    large species, generate_families)."""
    if name == "10sp":
        base = ["H2", "O2", "H2O", "OH", "H", "O", "HO2", "H2O2", "O3", "N2"]
        return generate(10, 27, seed if seed is not None else 1001, ("H", "O", "N"),
                        0.25, 0.15, inert=("N2",), majors=["H2", "O2", "H2O", "OH"],
                        base=base)
    if name == "53sp":
        # seed 1056: chosen by a scan of seeds 1000-1059 (64 reactors each) as
        # the mechanism whose cold mode (700-950 K) stays quiescent over dt
        # while combustible hot-mode (1050-1400 K) reactors ignite within dt
        return generate(53, 325, seed if seed is not None else 1056,
                        ("C", "H", "O", "N", "AR"), 0.25, 0.15, inert=("N2", "AR"),
                        majors=_C2_BASE, base=_C2_BASE)
    if name == "160sp":
        # seed 1089: chosen by a scan of seeds 1000-1120 (64-256 reactors each) as
        # a mechanism with no failed or endothermically collapsing reactor over
        # dt and a few percent of reactors igniting
        return generate(160, 900, seed if seed is not None else 1089,
                        ("C", "H", "O", "N", "AR"), 0.30, 0.10, inert=("N2", "AR"),
                        majors=_C2_BASE + ["C2H6", "CH3OH"], base=_C2_BASE)
    if name == "2000sp":
        return generate_families(2000, 8000, seed if seed is not None else 2001)
    raise ValueError(f"unknown workload {name}")


@dataclass
class States:
    T: np.ndarray
    P: np.ndarray
    Y: np.ndarray   # [n_reactors, n_species] reactor-major (AoS), as the C API
    seed: int


def workload_states(name: str, mech: Mechanism, n: int, seed: int) -> States:
    """Seeded reactor states of a named workload."""
    rng = np.random.default_rng(seed)
    nsp = len(mech.species)
    idx = {s: i for i, s in enumerate(mech.species)}
    Y = np.zeros((n, nsp))
    if name == "10sp":
        T = rng.uniform(900.0, 1500.0, n)
        P = rng.uniform(1.0, 20.0, n) * 101325.0
        maj = [idx[s] for s in mech.majors]
        frac = rng.uniform(0.3, 0.7, n)
        Y[:, maj] = rng.dirichlet(np.ones(len(maj)), n) * frac[:, None]
        Y[:, idx["N2"]] = 1.0 - frac
    elif name == "53sp":
        slow = rng.random(n) < 0.5
        T = np.where(slow, rng.uniform(700.0, 950.0, n), rng.uniform(1050.0, 1400.0, n))
        P = np.exp(rng.uniform(math.log(1.0), math.log(50.0), n)) * 101325.0
        maj = [idx[s] for s in mech.majors]
        Y[:, maj] = rng.dirichlet(0.5 * np.ones(len(maj)), n)
    elif name == "2000sp":
        T = rng.uniform(650.0, 1100.0, n)
        P = np.exp(rng.uniform(math.log(10.0), math.log(60.0), n)) * 101325.0
        maj = [idx[s] for s in mech.majors]
        Y[:, maj] = rng.dirichlet(0.3 * np.ones(len(maj)), n)
    elif name == "160sp":
        T = rng.uniform(800.0, 1600.0, n)
        P = np.exp(rng.uniform(math.log(1.0), math.log(40.0), n)) * 101325.0
        maj = [idx[s] for s in mech.majors]
        Y[:, maj] = rng.dirichlet(0.5 * np.ones(len(maj)), n)
    else:
        raise ValueError(name)
    # renormalise in extended precision so that sum(Y) = 1 to the last bit
    Y = Y / Y.sum(axis=1, keepdims=True)
    return States(T, P, np.ascontiguousarray(Y), seed)


def random_dense_states(mech: Mechanism, n: int, seed: int, Tlo=800.0, Thi=2200.0,
                        Plo=0.5, Phi=40.0) -> States:
    """States with every species present (rate/Jacobian parity cases)."""
    rng = np.random.default_rng(seed)
    nsp = len(mech.species)
    T = rng.uniform(Tlo, Thi, n)
    P = np.exp(rng.uniform(math.log(Plo), math.log(Phi), n)) * 101325.0
    Y = rng.dirichlet(np.ones(nsp), n)
    return States(T, P, Y, seed)


WORKLOAD_DT = {"10sp": 1.0e-4, "53sp": 1.0e-4, "160sp": 1.0e-5, "2000sp": 1.0e-4}
WORKLOAD_NR = {"10sp": 1024, "53sp": 65536, "160sp": 16384, "2000sp": 4096}
