"""Oracle-side CHEMKIN-II reader and reference table builder (test infrastructure).

An independent restatement, in Python/numpy, of what the reference does between
reading a mechanism and evaluating rates:

* CKReader text semantics (src/CKconverter/CKParser.cpp): sections
  ELEMENTS/SPECIES/THERMO/REACTIONS, '!' comments, 80-column NASA thermo cards
  (:861-1046), activation-energy units on the REACTIONS line (:1094-1110),
  auxiliary keywords LOW/TROE/SRI/REV/DUPLICATE/FORD/RORD/PLOG and third-body
  efficiencies (:1477-1669), '+M', '(+M)' and '(+species)' third bodies, PLOG
  exclusions (:312-345).
* Molecular weights: sum of element count x IUPAC-2007 atomic mass, in card
  order (src/zerork/species.cpp:38-45, atomicMassDB.h).
* Step numbering: forward step of reaction j, then its reverse step right after
  it when reversible with a non-zero reverse A (src/zerork/info_net.cpp:46-99);
  reactants/products expanded by multiplicity (:179-300).
* Rate-constant tables with the reference unit conversions
  (src/zerork/rate_const.cpp:63-81, 212-260): A*1000^(1-order), /1000 more for
  '+M' reactions, Ta = E*convertE, merged Arrhenius triples within 1e-6
  relative (:275-332, 1093-1128), Keq list (:334-386), third-body list with
  efficiency-1 in species-name order (:388-420, 1030-1062), falloff list with
  param[0] = ln(Alow*1000^-order) (:422-537), SRI parameters a, b, 1/c (-1 when
  c <= 0), d, e (:508-531).
* Non-integer reactions (info_net.cpp:411-470: FORD/RORD, or a coefficient
  further than 1e-14 from an integer): no expanded participants; per step
  rate-of-progress exponents by species name (summed coefficients, overwritten
  by FORD/RORD), reactant/product coefficient lists in text order, the step's
  real order = sum of its exponents for the unit conversions, Keq change in
  moles = reverse minus forward exponent sums
  (non_integer_reaction_network.cpp:53-262, rate_const.cpp:334-386).
* PLOG (rate_const.cpp:1146-1186, plog_reaction.cpp:10-135): pressures in Pa,
  A*1000^(1-order), selection sort by pressure, lines within 1e-8 relative of
  the previous distinct pressure grouped with it.

Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline use this.
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass, field

import numpy as np

NIST_RU = 8.314462618e3
CAL_PER_MOL_TACT = 4184.0 / NIST_RU
UNITS_TACT = {
    "CAL/MOLE": CAL_PER_MOL_TACT,
    "KCAL/MOLE": CAL_PER_MOL_TACT * 1000.0,
    "KELVINS": 1.0,
    "KJOULES/MOLE": 238.846 * CAL_PER_MOL_TACT,
    "JOULES/MOLE": 238.846 * CAL_PER_MOL_TACT / 1000.0,
}
# IUPAC 2007 atomic masses [kg/kmol] of elements 1-92, by symbol
# (src/zerork/atomicMassDB.h); looked up on the first two characters
ATOMIC_MASS = {
    "H": 1.00794, "HE": 4.002602, "LI": 6.941, "BE": 9.012182, "B": 10.811, "C": 12.0107,
    "N": 14.0067, "O": 15.9994, "F": 18.9984032, "NE": 20.1797, "NA": 22.98976928, "MG": 24.3050,
    "AL": 26.9815386, "SI": 28.0855, "P": 30.973762, "S": 32.065, "CL": 35.453, "AR": 39.948,
    "K": 39.0983, "CA": 40.078, "SC": 44.955912, "TI": 47.867, "V": 50.9415, "CR": 51.9961,
    "MN": 54.938045, "FE": 55.845, "CO": 58.933195, "NI": 58.6934, "CU": 63.546, "ZN": 65.38,
    "GA": 69.723, "GE": 72.64, "AS": 74.92160, "SE": 78.96, "BR": 79.904, "KR": 83.798,
    "RB": 85.4678, "SR": 87.62, "Y": 88.90585, "ZR": 91.224, "NB": 92.90638, "MO": 95.96,
    "TC": 98, "RU": 101.07, "RH": 102.90550, "PD": 106.42, "AG": 107.8682, "CD": 112.411,
    "IN": 114.818, "SN": 118.710, "SB": 121.760, "TE": 127.60, "I": 126.90447, "XE": 131.293,
    "CS": 132.9054519, "BA": 137.327, "LA": 138.90547, "CE": 140.116, "PR": 140.90765, "ND": 144.242,
    "PM": 145, "SM": 150.36, "EU": 151.964, "GD": 157.25, "TB": 158.92535, "DY": 162.500,
    "HO": 164.93032, "ER": 167.259, "TM": 168.93421, "YB": 173.054, "LU": 174.9668, "HF": 178.49,
    "TA": 180.94788, "W": 183.84, "RE": 186.207, "OS": 190.23, "IR": 192.217, "PT": 195.084,
    "AU": 196.966569, "HG": 200.59, "TL": 204.3833, "PB": 207.2, "BI": 208.98040, "PO": 209,
    "AT": 210, "RN": 222, "FR": 223, "RA": 226, "AC": 227, "TH": 232.03806,
    "PA": 231.03588, "U": 238.02891,
}

FALLOFF_LINDEMANN, FALLOFF_TROE3, FALLOFF_TROE4, FALLOFF_SRI = 0, 1, 2, 3
STOICH_TOL = 1.0e-14   # src/zerork/constants.h:32
P_ATM = 1.01325e5


@dataclass
class RxnText:
    reac: list           # [(name, coef)]
    prod: list
    reversible: bool
    A: float
    n: float
    E: float
    third_body: str = ""      # "" none, "M" +M, or "(+M)"/"(+X)" falloff species
    falloff: bool = False
    low: tuple = ()
    troe: tuple = ()
    rev: tuple = ()           # explicit reverse (A, n, E)
    eff: dict = field(default_factory=dict)
    dup: bool = False
    sri: tuple = ()
    plog: list = field(default_factory=list)   # [(p atm, A, n, E K)]
    ford: dict = field(default_factory=dict)
    rord: dict = field(default_factory=dict)
    real_order: bool = False


def _strip(line: str) -> str:
    i = line.find("!")
    return line if i < 0 else line[:i]


def _num(s: str) -> float:
    return float(s.replace("D", "E").replace("d", "e"))


def _sections(text: str):
    """Split a CHEMKIN file into [keyword, lines, rest-of-keyword-line] sections."""
    out = []
    cur = None
    for raw in text.splitlines():
        line = _strip(raw)
        tok = line.split()
        if not tok:
            continue
        key = tok[0].upper()
        if key == "END":
            cur = None
            continue
        if key[:4] in ("ELEM", "SPEC", "THER", "REAC") and not (cur and cur[0] == "THER" and len(raw) >= 80):
            cur = [key[:4], [], tok[1:]]
            out.append(cur)
            continue
        if cur is not None:
            cur[1].append(raw.rstrip("\n") if cur[0] == "THER" else line)
    return out


def _parse_thermo(lines, names_wanted):
    """80-column NASA cards -> {name: (tmid, low7, high7, [(el, n)])}."""
    res = {}
    i = 0
    # optional temperature line (three numbers)
    while i < len(lines) and not _strip(lines[i]).strip():
        i += 1
    tdef = [300.0, 1000.0, 5000.0]
    if i < len(lines):
        tok = _strip(lines[i]).split()
        try:
            vals = [_num(t) for t in tok[:3]]
            if len(vals) == 3 and (len(lines[i]) < 80 or lines[i][79] != "1"):
                tdef = vals
                i += 1
        except ValueError:
            pass
    while i < len(lines):
        s = lines[i]
        if len(s) < 80 or s[79] != "1":
            i += 1
            continue
        name = s[:16].split()[0]
        comp = []
        for k in range(4):
            loc = 24 + 5 * k
            sym = s[loc:loc + 2].strip()
            cnt = s[loc + 2:loc + 5].strip()
            if sym and cnt:
                n = int(round(_num(cnt)))
                if n != 0:
                    comp.append((sym.upper(), n))
        tmid_s = s[65:73].strip()
        tmid = _num(tmid_s) if tmid_s else tdef[1]
        if tmid == 0.0:
            tmid = tdef[1]
        if len(s) >= 78:
            sym = s[73:75].strip()
            cnt = s[75:78].strip()
            if sym and cnt:
                n = int(round(_num(cnt)))
                if n != 0:
                    comp.append((sym.upper(), n))
        l2, l3, l4 = lines[i + 1], lines[i + 2], lines[i + 3]
        c = [_num(l2[15 * k:15 * k + 15]) for k in range(5)]
        c += [_num(l3[15 * k:15 * k + 15]) for k in range(5)]
        c += [_num(l4[15 * k:15 * k + 15]) for k in range(4)]
        high, low = c[0:7], c[7:14]
        if name not in res:
            res[name] = (tmid, low, high, comp)
        i += 4
    return res


_SPLIT_EQ = re.compile(r"<=>|=>|=")


def _ln(a: float) -> float:
    """C's log: -inf for 0 (an A-factor of 0 gives a zero rate, as in
    rate_const.cpp's log of the A-factor), NaN for negative values."""
    if a > 0.0:
        return math.log(a)
    return -math.inf if a == 0.0 else math.nan


def _split_species(side: str):
    """'2H + O2' -> [('H',2), ('O2',1)]; '+M' handled by caller."""
    out = []
    # split on '+' that are separators (a '+' directly after another '+' is part of a name)
    parts = []
    cur = ""
    for ch in side:
        if ch == "+" and cur.strip():
            parts.append(cur)
            cur = ""
        else:
            cur += ch
    if cur.strip():
        parts.append(cur)
    for p in parts:
        p = p.strip()
        m = re.match(r"^(\d+(\.\d*)?)\s*(.+)$", p)
        if m and not p[0].isalpha():
            coef, name = float(m.group(1)), m.group(3).strip()
        else:
            coef, name = 1.0, p
        out.append((name, coef))
    return out


def _parse_reactions(lines, units_tok, species, ru=NIST_RU):
    cal = 4184.0 / ru
    conv = cal
    for t in units_tok:
        tu = t.upper()
        for k in ("KCAL/MOL", "KJOULES/MOL", "JOULES/MOL", "CAL/MOL", "KELVIN"):
            if tu.startswith(k):
                conv = {"KCAL/MOL": cal * 1000.0, "KJOULES/MOL": 238.846 * cal,
                        "JOULES/MOL": 238.846 * cal / 1000.0, "CAL/MOL": cal,
                        "KELVIN": 1.0}[k]
                break
    spset = set(species)
    rxns = []
    for line in lines:
        s = line.strip()
        if not s:
            continue
        if _SPLIT_EQ.search(s) and "/" not in s.split()[0]:
            tok = s.split()
            A, n, E = _num(tok[-3]), _num(tok[-2]), _num(tok[-1])
            eq = s[: s.rfind(tok[-3])]
            eq = eq[: len(eq)].rstrip()
            # remove the three trailing numbers robustly
            body = " ".join(tok[:-3])
            m = _SPLIT_EQ.search(body)
            sep = m.group(0)
            lhs, rhs = body[: m.start()], body[m.end():]
            rx = RxnText([], [], sep != "=>", A, n, E * conv)
            for side_i, side in enumerate((lhs, rhs)):
                fm = re.search(r"\(\s*\+\s*([^)\s]+)\s*\)", side)
                if fm:
                    rx.falloff = True
                    rx.third_body = fm.group(1)
                    side = side[: fm.start()] + side[fm.end():]
                sp = _split_species(side)
                keep = []
                for name, coef in sp:
                    if name.upper() == "M":
                        rx.third_body = "M"
                    else:
                        keep.append((name, coef))
                if side_i == 0:
                    rx.reac = keep
                else:
                    rx.prod = keep
            rxns.append(rx)
            continue
        # auxiliary line
        rx = rxns[-1]
        toks = re.findall(r"([^/\s]+)\s*(?:/([^/]*)/)?", s)
        for name, val in toks:
            key = name.upper()
            vals = val.split() if val else []
            if key in ("DUP", "DUPLICATE"):
                rx.dup = True
            elif key == "LOW":
                rx.low = (_num(vals[0]), _num(vals[1]), _num(vals[2]) * conv)
            elif key == "TROE":
                rx.troe = tuple(_num(v) for v in vals)
            elif key == "REV":
                rx.rev = (_num(vals[0]), _num(vals[1]), _num(vals[2]) * conv)
            elif key == "SRI":
                if rx.troe:
                    raise ValueError("cannot specify both SRI and TROE")
                rx.sri = tuple(_num(v) for v in vals)
            elif key == "PLOG":
                if len(vals) != 4:
                    raise ValueError("expected 4 PLOG parameters")
                rx.plog.append((_num(vals[0]), _num(vals[1]), _num(vals[2]), _num(vals[3]) * conv))
            elif key in ("FORD", "RORD"):
                (rx.ford if key == "FORD" else rx.rord)[vals[0]] = _num(vals[1])
                rx.real_order = True
            elif key in ("HIGH", "LT", "RLT", "CHEB", "TCHEB", "PCHEB"):
                raise NotImplementedError(f"oracle parser: keyword {key} not supported")
            else:
                if val is not None and vals:
                    rx.eff[name] = _num(vals[0])
    return rxns


@dataclass
class Tables:
    """Reference-layout mechanism tables (all FP64 / int32 numpy arrays)."""
    species: list
    mw: np.ndarray
    thermo: np.ndarray           # [nsp,16]: Tmid, low a1..a7, high a1..a7
    nrxn: int
    step_rxn: np.ndarray         # [nstep]
    step_dir: np.ndarray         # [nstep] +1 / -1
    reac_ptr: np.ndarray
    reac_sp: np.ndarray
    prod_ptr: np.ndarray
    prod_sp: np.ndarray
    arr_step: np.ndarray         # Arrhenius steps
    arr_lnA: np.ndarray
    arr_beta: np.ndarray
    arr_Ta: np.ndarray
    keq_step: np.ndarray
    keq_fwd: np.ndarray
    keq_ndelta: np.ndarray
    tb_fwd: np.ndarray
    tb_rev: np.ndarray
    tb_ptr: np.ndarray
    tb_sp: np.ndarray
    tb_eff: np.ndarray           # efficiency - 1
    fo_fwd: np.ndarray
    fo_rev: np.ndarray
    fo_type: np.ndarray
    fo_spc: np.ndarray           # -1 for M
    fo_param: np.ndarray         # [nfo,8]: lnAlow, beta, Ta, then Troe alpha, T***, T*, T**
                                 # or SRI a, b, 1/c (-1 if c <= 0), d, e
    fo_ptr: np.ndarray
    fo_sp: np.ndarray
    fo_eff: np.ndarray
    # non-integer steps (in step order)
    step_ni: np.ndarray          # [nstep] index into the ni lists or -1
    ni_step: np.ndarray
    ni_rop_ptr: np.ndarray       # CSR: rate-of-progress exponent species / powers
    ni_rop_sp: np.ndarray
    ni_rop_pow: np.ndarray
    ni_re_ptr: np.ndarray        # CSR: destroyed species / coefficients
    ni_re_sp: np.ndarray
    ni_re_nu: np.ndarray
    ni_pr_ptr: np.ndarray        # CSR: created species / coefficients
    ni_pr_sp: np.ndarray
    ni_pr_nu: np.ndarray
    # PLOG forward steps
    plog_step: np.ndarray
    plog_ptr: np.ndarray         # CSR over distinct pressures
    plog_p: np.ndarray           # Pa
    plog_lnp: np.ndarray
    plog_lptr: np.ndarray        # CSR (over all pressures) of Arrhenius lines
    plog_lnA: np.ndarray         # ln|A| with the unit conversion
    plog_sign: np.ndarray
    plog_beta: np.ndarray
    plog_Ta: np.ndarray

    @property
    def nsp(self):
        return len(self.species)

    @property
    def nstep(self):
        return len(self.step_rxn)


def build_tables(mech_path: str, therm_path: str | None, ru: float = NIST_RU) -> Tables:
    text = open(mech_path).read()
    secs = _sections(text)
    elements, species, rx_lines, units = [], [], [], []
    thermo_lines = []
    for key, lines, rest in secs:
        if key == "ELEM":
            for l in [" ".join(rest)] + lines:
                for t in l.split():
                    t = t.split("/")[0]
                    if t:
                        elements.append(t.upper())
        elif key == "SPEC":
            # a repeated name is ignored (CKParser.cpp:647-657)
            for l in [" ".join(rest)] + lines:
                for t in l.split():
                    if t not in species:
                        species.append(t)
        elif key == "THER":
            thermo_lines += lines
        elif key == "REAC":
            units = rest
            rx_lines += lines
    if therm_path:
        for key, lines, rest in _sections(open(therm_path).read()):
            if key == "THER":
                thermo_lines = lines + thermo_lines
    th = _parse_thermo(thermo_lines, species)
    thl = {k.upper(): v for k, v in th.items()}
    nsp = len(species)
    mw = np.zeros(nsp)
    thermo = np.zeros((nsp, 16))
    for j, s in enumerate(species):
        tmid, low, high, comp = th[s] if s in th else thl[s.upper()]
        w = 0.0
        for el, n in comp:
            w += n * ATOMIC_MASS[el.upper()[:2]]
        mw[j] = w
        thermo[j, 0] = tmid
        thermo[j, 1:8] = low
        thermo[j, 8:15] = high
    idx = {s: j for j, s in enumerate(species)}
    rxns = _parse_reactions(rx_lines, units, species, ru)
    nrxn = len(rxns)

    def is_nonint(r):
        if r.real_order:
            return True
        for _, c in r.reac + r.prod:
            if abs(c - int(c + 0.5)) > STOICH_TOL:
                return True
        return False

    for r in rxns:
        if r.plog:
            if r.falloff or r.low or r.troe or r.sri:
                raise ValueError("PLOG and falloff reactions (LOW/TROE/SRI) are mutually exclusive")
            if r.third_body == "M" and not r.falloff:
                raise ValueError("PLOG and 3rd-body reactions M or (M) are mutually exclusive")
            if r.rev and r.rev[0] > 0.0:
                raise ValueError("PLOG does not support REV parameters being specified")

    step_rxn, step_dir = [], []
    reac_lists, prod_lists = [], []
    step_order = []
    ni = []  # (step, rop [(sp, pow)], re [(sp, nu)], pr [(sp, nu)])

    def expand(lst):
        out = []
        for name, coef in lst:
            out += [idx[name]] * int(coef + 0.5)
        return out

    def rop_map(side, over):
        mp = {}
        for name, c in side:
            mp[name] = mp.get(name, 0.0) + c
        mp.update(over)
        return [(idx[n], mp[n]) for n in sorted(mp)]

    for j, r in enumerate(rxns):
        nonint = is_nonint(r)
        has_rev = r.reversible and not (r.rev and r.rev[0] == 0.0)
        re_l = [] if nonint else expand(r.reac)
        pr_l = [] if nonint else expand(r.prod)
        for d in range(2 if has_rev else 1):
            s_id = len(step_rxn)
            step_rxn.append(j); step_dir.append(1 if d == 0 else -1)
            reac_lists.append(re_l if d == 0 else pr_l)
            prod_lists.append(pr_l if d == 0 else re_l)
            if nonint:
                fwd_map, rev_map = rop_map(r.reac, r.ford), rop_map(r.prod, r.rord)
                rop = fwd_map if d == 0 else rev_map
                sre = [(idx[n], c) for n, c in (r.reac if d == 0 else r.prod)]
                spr = [(idx[n], c) for n, c in (r.prod if d == 0 else r.reac)]
                ni.append((s_id, rop, sre, spr))
                step_order.append(sum(p for _, p in rop))
            else:
                step_order.append(float(len(reac_lists[-1])))
    nstep = len(step_rxn)
    fwd_of = {}
    rev_of = {}
    for s_ in range(nstep):
        (fwd_of if step_dir[s_] == 1 else rev_of)[step_rxn[s_]] = s_
    step_ni = np.full(nstep, -1, dtype=np.int32)
    for q, e in enumerate(ni):
        step_ni[e[0]] = q

    def csr(lists):
        ptr = np.zeros(len(lists) + 1, dtype=np.int32)
        ptr[1:] = np.cumsum([len(l) for l in lists])
        return ptr, np.array([x for l in lists for x in l], dtype=np.int32)

    reac_ptr, reac_sp = csr(reac_lists)
    prod_ptr, prod_sp = csr(prod_lists)

    # Arrhenius steps with the reference unit conversion, then merging; PLOG
    # forward steps have their own tables
    arr = []  # (A, Tpow, Tact, step)
    keq = []
    plog = []  # (step, [(p, lnA, sign, beta, Ta) sorted], groups)
    for j, r in enumerate(rxns):
        s_ = fwd_of[j]
        tb = r.third_body == "M" and not r.falloff
        if not r.plog:
            A = r.A * 1000.0 ** (1.0 - step_order[s_])
            if tb:
                A /= 1000.0
            arr.append((A, r.n, r.E, s_))
        else:
            ac = 1000.0 ** (1.0 - step_order[s_])
            ln = [[p * P_ATM, A * ac, n, E] for p, A, n, E in r.plog]
            for a in range(len(ln) - 1):      # selection sort (plog_reaction.cpp:54-94)
                mi = a
                for b in range(a + 1, len(ln)):
                    if ln[b][0] < ln[mi][0]:
                        mi = b
                if mi != a:
                    ln[a], ln[mi] = ln[mi], ln[a]
            pts, lptr = [], []
            for a, l in enumerate(ln):
                if not pts or abs(l[0] - pts[-1]) > abs(pts[-1] * 1.0e-8):
                    pts.append(l[0])
                    lptr.append(a)
            lptr.append(len(ln))
            plog.append((s_, pts, lptr, ln))
        if j in rev_of:
            sr = rev_of[j]
            if r.rev and r.rev[0] > 0.0:
                Ar = r.rev[0] * 1000.0 ** (1.0 - step_order[sr])
                if tb:
                    Ar /= 1000.0
                arr.append((Ar, r.rev[1], r.rev[2], sr))
            else:
                if step_ni[s_] >= 0:
                    nd = step_order[sr] - step_order[s_]
                else:
                    nd = float(len(prod_lists[s_]) - len(reac_lists[s_]))
                keq.append((sr, s_, nd))
    arr_sorted = sorted(arr, key=lambda e: (e[0], -e[2], e[1]))

    def same(x, y):
        for a, b in ((x[0], y[0]), (x[1], y[1]), (x[2], y[2])):
            if abs(a - b) > 1.0e-6 * abs(0.5 * (a + b)) and abs(a - b) > 0.0:
                return False
        return True

    rep = []
    last = None
    for e in arr_sorted:
        if last is None or not same(last, e):
            last = e
        rep.append((e[3], _ln(last[0]), last[1], last[2]))
    rep.sort(key=lambda t: t[0])
    arr_step = np.array([t[0] for t in rep], dtype=np.int32)
    arr_lnA = np.array([t[1] for t in rep])
    arr_beta = np.array([t[2] for t in rep])
    arr_Ta = np.array([t[3] for t in rep])
    keq_step = np.array([k[0] for k in keq], dtype=np.int32)
    keq_fwd = np.array([k[1] for k in keq], dtype=np.int32)
    keq_nd = np.array([k[2] for k in keq])

    def eff_list(r):
        out = []
        for name in sorted(r.eff.keys()):
            if name in idx:
                out.append((idx[name], r.eff[name] - 1.0))
        return out

    tb_fwd, tb_rev, tb_l = [], [], []
    fo_fwd, fo_rev, fo_type, fo_spc, fo_par, fo_l = [], [], [], [], [], []
    for j, r in enumerate(rxns):
        if r.third_body == "M" and not r.falloff:
            tb_fwd.append(fwd_of[j]); tb_rev.append(rev_of.get(j, -1)); tb_l.append(eff_list(r))
        if r.falloff:
            s_ = fwd_of[j]
            fo_fwd.append(s_); fo_rev.append(rev_of.get(j, -1))
            single = r.third_body.upper() != "M"
            fo_spc.append(idx[r.third_body] if single else -1)
            fo_l.append(eff_list(r) if not single else [])
            p = [_ln(r.low[0] * 1000.0 ** (-step_order[s_])), r.low[1], r.low[2],
                 0.0, 0.0, 0.0, 0.0, 0.0]
            if r.sri:
                c = r.sri[2]
                p[3:8] = [r.sri[0], r.sri[1], 1.0 / c if c > 0 else -1.0,
                          r.sri[3] if len(r.sri) >= 4 else 1.0,
                          r.sri[4] if len(r.sri) == 5 else 0.0]
                fo_type.append(FALLOFF_SRI)
            elif r.troe:
                p[3:3 + len(r.troe)] = list(r.troe)
                fo_type.append(FALLOFF_TROE3 if len(r.troe) == 3 else FALLOFF_TROE4)
            else:
                fo_type.append(FALLOFF_LINDEMANN)
            fo_par.append(p)
    tb_ptr, tb_sp = csr([[e[0] for e in l] for l in tb_l])
    tb_eff = np.array([e[1] for l in tb_l for e in l])
    fo_ptr, fo_sp = csr([[e[0] for e in l] for l in fo_l])
    fo_eff = np.array([e[1] for l in fo_l for e in l])
    I = lambda v: np.array(v, dtype=np.int32)
    F = lambda v: np.array(v, dtype=np.float64)
    ni_rop_ptr, ni_rop_sp = csr([[a for a, _ in e[1]] for e in ni])
    ni_re_ptr, ni_re_sp = csr([[a for a, _ in e[2]] for e in ni])
    ni_pr_ptr, ni_pr_sp = csr([[a for a, _ in e[3]] for e in ni])
    plog_ptr = np.zeros(len(plog) + 1, dtype=np.int32)
    plog_ptr[1:] = np.cumsum([len(e[1]) for e in plog])
    lptr, lines = [0], []
    for e in plog:
        for k in range(len(e[1])):
            lines += e[3][e[2][k]:e[2][k + 1]]
            lptr.append(len(lines))
    return Tables(species, mw, thermo, nrxn, I(step_rxn), I(step_dir), reac_ptr, reac_sp,
                  prod_ptr, prod_sp, arr_step, arr_lnA, arr_beta, arr_Ta, keq_step, keq_fwd,
                  keq_nd, I(tb_fwd), I(tb_rev), tb_ptr, tb_sp, tb_eff, I(fo_fwd), I(fo_rev),
                  I(fo_type), I(fo_spc), F(fo_par).reshape(-1, 8), fo_ptr, fo_sp, fo_eff,
                  step_ni, I([e[0] for e in ni]), ni_rop_ptr, ni_rop_sp,
                  F([p for e in ni for _, p in e[1]]), ni_re_ptr, ni_re_sp,
                  F([c for e in ni for _, c in e[2]]), ni_pr_ptr, ni_pr_sp,
                  F([c for e in ni for _, c in e[3]]),
                  I([e[0] for e in plog]), plog_ptr, F([p for e in plog for p in e[1]]),
                  F([math.log(p) for e in plog for p in e[1]]), I(lptr),
                  F([_ln(abs(l[1])) for l in lines]),
                  F([-1.0 if l[1] < 0.0 else 1.0 for l in lines]),
                  F([l[2] for l in lines]), F([l[3] for l in lines]))
