"""Mechanism front-end: Cantera-YAML reaction mechanism -> dense intermediate representation.

Follows the reference generator's reading of a mechanism:
  * units check and unit conversion of rate constants (kinetix/core/mechanism.py:15-26,
    kinetix/core/reaction_rates.py:67-99);
  * reaction equation parsing, reaction type decoding, efficiencies
    (reaction_rates.py:30-190);
  * species order: YAML definition order with inert species (zero net stoichiometry in
    every reaction) moved last (kinetix/core/generate.py:36-47);
  * NASA-7 pieces (single-piece species duplicated with a 1000 K split) and transport
    data unit conversions (kinetix/core/mix_transport.py:210-251).

The IR is plain numpy arrays so that the HIP emitter, the CPU oracle and the tests all
consume the same tables.
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np
import yaml

from . import physics as phys

# Reaction kinds (also the integer codes stored in the IR).
ELEMENTARY, THREE_BODY, LINDEMANN, TROE, SRI, PLOG = range(6)
KIND_NAMES = ('elementary', 'three-body', 'lindemann', 'troe', 'sri', 'plog')


class _Loader(yaml.SafeLoader):
    """SafeLoader with YAML-1.2 scalars: ``1e+14`` is a float and only true/false are
    booleans (YAML 1.1 would read the species name ``NO`` as False)."""


_Loader.yaml_implicit_resolvers = {
    ch: [(tag, rx) for tag, rx in lst if tag != 'tag:yaml.org,2002:bool']
    for ch, lst in yaml.SafeLoader.yaml_implicit_resolvers.items()}
_Loader.add_implicit_resolver(
    'tag:yaml.org,2002:bool', re.compile(r'^(?:true|True|TRUE|false|False|FALSE)$'),
    list('tTfF'))
_Loader.add_implicit_resolver(
    'tag:yaml.org,2002:float',
    re.compile(r'''^(?:[-+]?(?:[0-9][0-9_]*)\.[0-9_]*(?:[eE][-+]?[0-9]+)?
                    |[-+]?(?:[0-9][0-9_]*)(?:[eE][-+]?[0-9]+)
                    |\.[0-9_]+(?:[eE][-+]?[0-9]+)?
                    |[-+]?\.(?:inf|Inf|INF)
                    |\.(?:nan|NaN|NAN))$''', re.X),
    list('-+0123456789.'))


def load_yaml(path):
    with open(path) as fh:
        return yaml.load(fh, Loader=_Loader)


@dataclass
class Reaction:
    equation: str
    nu_f: np.ndarray                 # reactant stoichiometric coefficients (N,)
    nu_r: np.ndarray                 # product stoichiometric coefficients (N,)
    kind: int
    reversible: bool
    A: float                         # k (or k_inf) pre-exponential, SI (m^3, mol, s)
    beta: float
    Ta: float                        # activation temperature E/R [K]
    A0: float = 0.0                  # low-pressure limit (falloff only)
    beta0: float = 0.0
    Ta0: float = 0.0
    efficiencies: Optional[np.ndarray] = None   # (N,) or None
    collider: int = -1               # explicit collider species index, -1 for M
    troe: Optional[tuple] = None     # (A, T3, T1, T2); T2 = inf when absent
    sri: Optional[tuple] = None      # (a, b, c, d, e)
    plog: Optional[list] = None      # [(P [Pa], [(A, beta, Ta), ...]), ...] sorted by P

    @property
    def net(self):
        return self.nu_r - self.nu_f

    @property
    def has_falloff(self):
        return self.kind in (LINDEMANN, TROE, SRI)


@dataclass
class Mechanism:
    name: str
    species: List[str]
    n_active: int
    W: np.ndarray                    # molar mass [kg/mol] (N,)
    elements: List[str]
    composition: np.ndarray          # (n_elements, N) atom counts
    nasa_lo: np.ndarray              # (N, 7) coefficients below T_mid
    nasa_hi: np.ndarray              # (N, 7) coefficients above T_mid
    T_mid: np.ndarray                # (N,)
    T_range: np.ndarray              # (N, 2) validity range [T_low, T_high]
    geometry_dof: np.ndarray         # rotational degrees of freedom: 0, 1 or 3/2
    well_depth: np.ndarray           # epsilon [J]
    diameter: np.ndarray             # sigma [m]
    dipole: np.ndarray               # [C m]
    polarizability: np.ndarray       # [m^3]
    rot_relax: np.ndarray            # rotational relaxation collision number at 298 K
    reactions: List[Reaction] = field(default_factory=list)

    @property
    def n_species(self):
        return len(self.species)

    @property
    def n_reactions(self):
        return len(self.reactions)

    def index(self, name):
        return self.species.index(name)


_FALLOFF_COLLIDER = re.compile(r'\s*\(\+\s*([^()]+(?:\([^()]*\))?)\s*\)')


def _split_equation(eq):
    """Return (lhs, rhs, reversible) of a reaction equation string."""
    if '<=>' in eq:
        lhs, rhs = eq.split('<=>')
        return lhs, rhs, True
    m = re.search(r'(?<!<)=>', eq)
    if m:
        return eq[:m.start()], eq[m.end():], False
    if '=' in eq:
        lhs, rhs = eq.split('=', 1)
        return lhs, rhs, True
    raise ValueError(f'cannot parse reaction equation {eq!r}')


def _parse_side(side, names):
    """Stoichiometric vector of one side; returns (nu, collider_name or None)."""
    nu = np.zeros(len(names), dtype=np.int64)
    collider = None
    m = _FALLOFF_COLLIDER.search(side)
    if m:
        collider = m.group(1).strip()
        side = side[:m.start()] + side[m.end():]
    for tok in side.split(' + '):
        tok = tok.strip()
        if not tok:
            continue
        parts = tok.split()
        if len(parts) == 2:
            coef, sp = parts
            c = float(coef)
            if c != int(c):
                raise ValueError(f'non-integer stoichiometric coefficient in {side!r}')
            coef = int(c)
        else:
            coef, sp = 1, parts[0]
        if sp == 'M':
            collider = collider or 'M'
            continue
        if sp not in names:
            raise ValueError(f'unknown species {sp!r} in {side!r}')
        nu[names.index(sp)] += coef
    return nu, collider


def _activation_temperature(Ea, units):
    unit = units.get('activation-energy', 'cal/mol')
    if unit == 'K':
        return Ea
    if unit == 'cal/mol':
        return Ea * phys.CAL / phys.R_GAS
    if unit == 'kcal/mol':
        return Ea * 1e3 * phys.CAL / phys.R_GAS
    if unit == 'J/mol':
        return Ea / phys.R_GAS
    if unit == 'kJ/mol':
        return Ea * 1e3 / phys.R_GAS
    raise ValueError(f'unknown activation-energy units {unit!r}')


def _rate(rc, units, conc_order):
    """(A_SI, beta, Ta) of a YAML rate constant; A scaled by (cm^3 -> m^3)^conc_order."""
    return (rc['A'] * pow(1e-6, conc_order), rc['b'], _activation_temperature(rc['Ea'], units))


def parse_reaction(r, names, units):
    eq = r['equation']
    lhs, rhs, reversible = _split_equation(eq)
    nu_f, coll_f = _parse_side(lhs, names)
    nu_r, _ = _parse_side(rhs, names)
    if 'reversible' in r and not r['reversible']:
        reversible = False
    rtype = r.get('type', 'elementary')
    order = int(nu_f.sum())
    collider = -1
    if coll_f is not None and coll_f != 'M' and coll_f in names:
        collider = names.index(coll_f)

    kw = {}
    if rtype == 'elementary':
        kind = ELEMENTARY
        A, b, Ta = _rate(r['rate-constant'], units, order - 1)
    elif rtype == 'three-body':
        kind = THREE_BODY
        A, b, Ta = _rate(r['rate-constant'], units, order)       # +1 order for [M]
    elif rtype == 'falloff':
        A, b, Ta = _rate(r['high-P-rate-constant'], units, order - 1)
        kw['A0'], kw['beta0'], kw['Ta0'] = _rate(r['low-P-rate-constant'], units, order)
        if r.get('Troe'):
            kind = TROE
            t = r['Troe']
            kw['troe'] = (float(t['A']), float(t['T3']), float(t['T1']), float(t.get('T2', math.inf)))
        elif r.get('SRI'):
            kind = SRI
            s = r['SRI']
            kw['sri'] = (float(s['A']), float(s['B']), float(s['C']), float(s.get('D', 1.0)),
                         float(s.get('E', 0.0)))
        else:
            kind = LINDEMANN
    elif rtype == 'pressure-dependent-Arrhenius':
        kind = PLOG
        table = {}
        for rc in r['rate-constants']:
            P = rc['P']
            if isinstance(P, str):
                val, unit = P.split()
                P = float(val) * {'atm': phys.ONE_ATM, 'Pa': 1.0, 'bar': 1e5}[unit]
            table.setdefault(float(P), []).append(_rate(rc, units, order - 1))
        kw['plog'] = sorted(table.items())
        A, b, Ta = kw['plog'][0][1][0]
    else:
        raise ValueError(f'unsupported reaction type {rtype!r} in {eq!r}')

    eff = None
    if r.get('efficiencies') is not None:
        default = r.get('default-efficiency', 1.0)
        eff = np.array([float(r['efficiencies'].get(s, default)) for s in names])
    return Reaction(eq, nu_f, nu_r, kind, reversible, A, b, Ta, efficiencies=eff,
                    collider=collider, **kw)


def from_model(model, name='mech'):
    units = model.get('units', {})
    if units.get('length', 'm') != 'cm' or units.get('quantity', 'kmol') != 'mol' \
            or units.get('time', 's') != 's':
        raise ValueError('mechanism units must be cm, mol, s')
    defs = model['species']
    names0 = [s['name'] for s in defs]
    rx0 = [parse_reaction(r, names0, units) for r in model['reactions']]
    net = np.array([rx.net for rx in rx0]) if rx0 else np.zeros((0, len(names0)), int)
    active = [i for i in range(len(defs)) if np.any(net[:, i] != 0)]
    inert = [i for i in range(len(defs)) if not np.any(net[:, i] != 0)]
    order = active + inert
    sp = [defs[i] for i in order]
    names = [s['name'] for s in sp]
    reactions = [parse_reaction(r, names, units) for r in model['reactions']]

    elements = sorted({e for s in sp for e in s['composition']},
                      key=lambda e: phys.ATOMIC_WEIGHT.get(e, 0.0))
    comp = np.array([[s['composition'].get(e, 0) for s in sp] for e in elements], dtype=float)
    W = np.array([sum(cnt * phys.ATOMIC_WEIGHT[el] / 1e3 for el, cnt in s['composition'].items())
                  for s in sp])

    lo, hi, tmid, trng = [], [], [], []
    for s in sp:
        th = s['thermo']
        if th.get('model', 'NASA7') != 'NASA7':
            raise ValueError(f"species {s['name']}: only NASA7 thermo is supported")
        data = [list(map(float, d)) for d in th['data']]
        tr = th['temperature-ranges']
        if len(data) > 2:
            raise ValueError(f"species {s['name']} has more than 2 NASA-7 pieces")
        if len(data) == 1:
            data = [data[0], list(data[0])]
            split = 1000.0
        else:
            split = float(tr[1])
        lo.append(data[0]); hi.append(data[1]); tmid.append(split)
        trng.append((float(tr[0]), float(tr[-1])))

    dof_map = {'atom': 0.0, 'linear': 1.0, 'nonlinear': 1.5}
    tp = [s.get('transport', {}) for s in sp]
    return Mechanism(
        name=name, species=names, n_active=len(active), W=W, elements=elements,
        composition=comp, nasa_lo=np.array(lo), nasa_hi=np.array(hi), T_mid=np.array(tmid),
        T_range=np.array(trng),
        geometry_dof=np.array([dof_map.get(t.get('geometry', 'atom'), 0.0) for t in tp]),
        well_depth=np.array([t.get('well-depth', 0.0) * phys.K_BOLTZMANN for t in tp]),
        diameter=np.array([t.get('diameter', 0.0) * 1e-10 for t in tp]),
        dipole=np.array([t.get('dipole', 0.0) * phys.DEBYE for t in tp]),
        polarizability=np.array([t.get('polarizability', 0.0) * 1e-30 for t in tp]),
        rot_relax=np.array([float(t.get('rotational-relaxation', 0.0)) for t in tp]),
        reactions=reactions)


def load(path, name=None):
    import os
    model = load_yaml(path)
    return from_model(model, name or os.path.splitext(os.path.basename(path))[0])
