"""Seeded synthetic mechanism: N species / R reversible reactions in Cantera YAML.

The species are drawn (with replacement, by seed) from a source mechanism's species set
(GRI-3.0): each synthetic species S### copies the composition, NASA-7 polynomials and
Lennard-Jones / dipole / polarizability data of its source species. Reactions:
  * 1-3 reactants and 1-3 products, drawn uniformly from the synthetic species;
  * A ~ logU[1e6, 1e14] (cm, mol, s units of the reaction's order), beta ~ U[-2, 2],
    Ea ~ U[0, 300] kJ/mol (written in cal/mol);
  * 15 % three-body (efficiencies ~ U[0.5, 5] on 5 random species), 10 % Troe falloff
    (k_0 ~ logU[1e12, 1e20], A ~ U[0.1, 0.9], T3 ~ logU[30, 1000], T1 ~ logU[800, 10000],
    T2 ~ logU[1000, 10000]), the rest elementary; all reversible.
Floats are written with a decimal point (the reference's loader needs dotted floats) and
Ea in cal/mol (kinetix/core/reaction_rates.py:84-89 accepts cal/mol or K only).

coverage=True writes every reaction form the generator emits instead (a small mechanism that
exercises the emitter paths no real straight-line mechanism here has): reversible and
irreversible elementary, three-body, Lindemann, Troe with and without T2, SRI with 3 and 5
parameters, falloff with an explicit collider species, and PLOG over 2-4 of {0.1, 1, 10,
100} atm, a quarter of them with two summed rates at one pressure.

    python -m kinetix_amd.synth --source gri30.yaml --species 160 --reactions 1000 \
        --seed 20261205 --out kinetix_amd/mechs/synth160/synth160.yaml
"""
from __future__ import annotations

import argparse
import os

import numpy as np

from .gen import mechanism

SEED = 20261202 + 3                    # the synthetic workload's seed (base + config index)


def _f(x):
    """Float with a decimal point or exponent (YAML-1.2 float, never an int)."""
    s = repr(float(x))
    return s if ('.' in s or 'e' in s or 'inf' in s or 'nan' in s) else s + '.0'


def _species_block(name, src):
    th = src['thermo']
    lines = [f'- name: {name}',
             '  composition: {' + ', '.join(f'{e}: {int(n)}' for e, n in src['composition'].items()) + '}',
             '  thermo:', '    model: NASA7',
             '    temperature-ranges: [' + ', '.join(_f(t) for t in th['temperature-ranges']) + ']',
             '    data:']
    for d in th['data']:
        lines.append('    - [' + ', '.join(_f(v) for v in d) + ']')
    tr = src.get('transport')
    if tr:
        lines += ['  transport:', '    model: gas', f"    geometry: {tr.get('geometry', 'atom')}"]
        for key in ('well-depth', 'diameter', 'dipole', 'polarizability', 'rotational-relaxation'):
            if key in tr:
                lines.append(f'    {key}: {_f(tr[key])}')
    return lines


def _side(names, stoich):
    return ' + '.join((f'{c} {n}' if c > 1 else n) for n, c in stoich.items())


def _arrh(rng, lo=6, hi=14):
    return (f'A: {_f(10 ** rng.uniform(lo, hi))}, b: {_f(rng.uniform(-2, 2))}, '
            f'Ea: {_f(rng.uniform(0, 300e3) / 4.184)}')


def _coverage_reaction(rng, names, reac, prod, kind):
    """YAML lines of one reaction of the coverage mechanism (kind 0..9)."""
    lhs, rhs = _side(names, reac), _side(names, prod)
    eff = lambda: ('  efficiencies: {' + ', '.join(
        f'{names[int(i)]}: {_f(rng.uniform(0.5, 5))}' for i in rng.choice(len(names), 5, replace=False)) + '}')
    low = lambda: f'  low-P-rate-constant: {{{_arrh(rng, 12, 20)}}}'
    high = lambda: f'  high-P-rate-constant: {{{_arrh(rng)}}}'
    if kind == 0:
        return [f'- equation: {lhs} <=> {rhs}', f'  rate-constant: {{{_arrh(rng)}}}']
    if kind == 1:
        return [f'- equation: {lhs} => {rhs}', f'  rate-constant: {{{_arrh(rng)}}}']
    if kind == 2:
        return [f'- equation: {lhs} + M <=> {rhs} + M', '  type: three-body',
                f'  rate-constant: {{{_arrh(rng)}}}', eff()]
    if kind in (3, 4, 5, 6, 7):
        out = [f'- equation: {lhs} (+M) <=> {rhs} (+M)', '  type: falloff', low(), high()]
        if kind == 4 or kind == 5:
            troe = (f'A: {_f(rng.uniform(0.1, 0.9))}, T3: {_f(10 ** rng.uniform(np.log10(30), 3))}, '
                    f'T1: {_f(10 ** rng.uniform(np.log10(800), 4))}')
            if kind == 4:
                troe += f', T2: {_f(10 ** rng.uniform(3, 4))}'
            out.append(f'  Troe: {{{troe}}}')
        elif kind in (6, 7):
            sri = (f'A: {_f(rng.uniform(0.1, 50))}, B: {_f(rng.uniform(-1000, 20000))}, '
                   f'C: {_f(rng.uniform(100, 5000))}')
            if kind == 7:
                sri += f', D: {_f(rng.uniform(0.5, 2))}, E: {_f(rng.uniform(-0.5, 0.5))}'
            out.append(f'  SRI: {{{sri}}}')
        return out + [eff()]
    if kind == 8:
        used = set(reac) | set(prod)
        coll = next(n for n in rng.permutation(names) if n not in used)
        return [f'- equation: {lhs} (+{coll}) <=> {rhs} (+{coll})', '  type: falloff', low(), high(),
                f'  Troe: {{A: {_f(rng.uniform(0.1, 0.9))}, T3: {_f(10 ** rng.uniform(np.log10(30), 3))}, '
                f'T1: {_f(10 ** rng.uniform(np.log10(800), 4))}}}']
    pts = sorted(rng.choice([0.1, 1.0, 10.0, 100.0], size=int(rng.integers(2, 5)), replace=False))
    out = [f'- equation: {lhs} <=> {rhs}', '  type: pressure-dependent-Arrhenius', '  rate-constants:']
    for i, P in enumerate(pts):
        for _ in range(2 if (i == 0 and rng.uniform() < 0.25) else 1):
            out.append(f'  - {{P: {_f(P)} atm, {_arrh(rng)}}}')
    return out


def write(source_yaml, n_species, n_reactions, seed, out_path, coverage=False):
    src = mechanism.load_yaml(source_yaml)
    pool = [s for s in src['species']]
    rng = np.random.default_rng(seed)
    pick = rng.integers(0, len(pool), size=n_species)
    names = [f'S{k:03d}' for k in range(n_species)]
    lines = ['description: seeded synthetic mechanism (species resampled from '
             f'{os.path.basename(source_yaml)}, seed {seed})',
             'units: {length: cm, time: s, quantity: mol, activation-energy: cal/mol}',
             'phases:', f'- name: synth{n_species}', '  thermo: ideal-gas',
             '  elements: [' + ', '.join(sorted({e for i in pick for e in pool[i]['composition']})) + ']',
             '  species: [' + ', '.join(names) + ']', '  kinetics: gas',
             '  transport: mixture-averaged', '  state: {T: 300.0, P: 1 atm}', '', 'species:']
    for k, i in enumerate(pick):
        lines += _species_block(names[k], pool[i])
    lines += ['', 'reactions:']
    if coverage:
        kinds = rng.choice(10, size=n_reactions, p=[0.3, 0.1, 0.12, 0.08, 0.08, 0.06, 0.06, 0.06, 0.06, 0.08])
        for r in range(n_reactions):
            nr, npd = int(rng.integers(1, 4)), int(rng.integers(1, 4))
            reac, prod = {}, {}
            for _ in range(nr):
                sp = names[int(rng.integers(0, n_species))]
                reac[sp] = reac.get(sp, 0) + 1
            for _ in range(npd):
                sp = names[int(rng.integers(0, n_species))]
                prod[sp] = prod.get(sp, 0) + 1
            lines += _coverage_reaction(rng, names, reac, prod, int(kinds[r]))
            lines.append('  duplicate: true')
        os.makedirs(os.path.dirname(os.path.abspath(out_path)), exist_ok=True)
        with open(out_path, 'w') as fh:
            fh.write('\n'.join(lines) + '\n')
        return out_path
    kind = rng.choice(3, size=n_reactions, p=[0.75, 0.15, 0.10])     # elementary, 3-body, Troe
    for r in range(n_reactions):
        nr, npd = int(rng.integers(1, 4)), int(rng.integers(1, 4))
        reac, prod = {}, {}
        for _ in range(nr):
            s = names[int(rng.integers(0, n_species))]
            reac[s] = reac.get(s, 0) + 1
        for _ in range(npd):
            s = names[int(rng.integers(0, n_species))]
            prod[s] = prod.get(s, 0) + 1
        A = 10 ** rng.uniform(6, 14)
        b = rng.uniform(-2, 2)
        Ea = rng.uniform(0, 300e3) / 4.184                       # J/mol -> cal/mol
        rate = f'{{A: {_f(A)}, b: {_f(b)}, Ea: {_f(Ea)}}}'
        if kind[r] == 0:
            lines += [f'- equation: {_side(names, reac)} <=> {_side(names, prod)}',
                      f'  rate-constant: {rate}']
        elif kind[r] == 1:
            lines += [f'- equation: {_side(names, reac)} + M <=> {_side(names, prod)} + M',
                      '  type: three-body', f'  rate-constant: {rate}']
        else:
            A0 = 10 ** rng.uniform(12, 20)
            b0 = rng.uniform(-2, 2)
            E0 = rng.uniform(0, 300e3) / 4.184
            lines += [f'- equation: {_side(names, reac)} (+M) <=> {_side(names, prod)} (+M)',
                      '  type: falloff',
                      f'  low-P-rate-constant: {{A: {_f(A0)}, b: {_f(b0)}, Ea: {_f(E0)}}}',
                      f'  high-P-rate-constant: {rate}',
                      f'  Troe: {{A: {_f(rng.uniform(0.1, 0.9))}, T3: {_f(10 ** rng.uniform(np.log10(30), 3))}, '
                      f'T1: {_f(10 ** rng.uniform(np.log10(800), 4))}, T2: {_f(10 ** rng.uniform(3, 4))}}}']
        if kind[r] != 0:
            eff = {names[int(i)]: rng.uniform(0.5, 5) for i in rng.choice(n_species, 5, replace=False)}
            lines.append('  efficiencies: {' + ', '.join(f'{k}: {_f(v)}' for k, v in eff.items()) + '}')
        lines.append('  duplicate: true')
    os.makedirs(os.path.dirname(os.path.abspath(out_path)), exist_ok=True)
    with open(out_path, 'w') as fh:
        fh.write('\n'.join(lines) + '\n')
    return out_path


def main(argv=None):
    ap = argparse.ArgumentParser(description='write a seeded synthetic Cantera-YAML mechanism')
    ap.add_argument('--source', required=True, help='source mechanism (species data)')
    ap.add_argument('--species', type=int, default=160)
    ap.add_argument('--reactions', type=int, default=1000)
    ap.add_argument('--seed', type=int, default=SEED)
    ap.add_argument('--out', required=True)
    ap.add_argument('--coverage', action='store_true', help='every reaction form (see module doc)')
    a = ap.parse_args(argv)
    print(write(a.source, a.species, a.reactions, a.seed, a.out, coverage=a.coverage))


if __name__ == '__main__':
    main()
