"""Write the golden fixtures under tests/golden/ (run in the build container only).

Sources of truth:
  * gri30_ref.npz: seeded GRI-3.0 states and the outputs of the reference's OWN compiled
    FP64 routines (examples/KinetiX/mechanisms/gri/*, built into oracle/_ref by
    oracle/Makefile), with the reference SERIAL flags ("fast") and strict IEEE ("strict").
    The fast-vs-strict spread is the reference's own noise floor.
  * ci_data/gri30.*.cantera: the reference's Cantera known-answer records
    (kinetix/ci_data/), copied verbatim as data.

Usage: python tests/golden/make_golden.py   (needs /root/reference)
"""
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from kinetix_amd import states  # noqa: E402
from kinetix_amd.gen import tables  # noqa: E402
from oracle import pyoracle as po  # noqa: E402

REF = os.environ.get('REF', '/root/reference')


def edge_states(species, W):
    """Hand-picked edge cases: NASA-7 switch temperatures, fit-window ends, pure species,
    zero and negative mass fractions (clipped by the kernels), a near-pure N2 state."""
    N = len(species)
    cols = []
    air = states.air_mass_fractions(species, W)
    for T in (300.0, 1000.0, 1368.0, 1382.0, 1478.0, 2999.0, 3000.0):
        cols.append(np.concatenate([[T], air]))
    for name in ('H2', 'O2', 'H2O', 'N2', 'CH4', 'OH', 'AR'):
        Y = np.zeros(N)
        Y[species.index(name)] = 1.0
        cols.append(np.concatenate([[1500.0], Y]))
    Y = air.copy()
    Y[species.index('H')] = -1e-12          # tiny negative: clipped to 0 in the kernels
    cols.append(np.concatenate([[1800.0], Y]))
    Y = np.full(N, 1e-30)
    Y[species.index('N2')] = 1.0
    cols.append(np.concatenate([[2200.0], Y / Y.sum()]))
    return np.array(cols).T


def main():
    po.build_libs(with_ref=True)
    info, t = tables.load(os.path.join(ROOT, 'kinetix_amd', 'mechs', 'gri30'))
    sp, W = info['species'], t['W']
    ref_fast, ref_strict = po.Reference('fast'), po.Reference('strict')
    assert ref_fast.species == sp, 'species order differs from the reference build'

    s_bk1, _ = states.sample(384, sp, W, T_range=(1000.0, 2500.0), seed=states.BASE_SEED + 0)
    s_bk2, p_bk2 = states.sample(384, sp, W, T_range=(300.0, 2500.0), seed=states.BASE_SEED + 1,
                                 p_range_atm=(0.5, 40.0))
    edge = edge_states(sp, W)
    s_bk1 = np.concatenate([s_bk1, edge], axis=1)
    s_bk2 = np.concatenate([s_bk2, edge], axis=1)
    p_bk2 = np.concatenate([p_bk2, np.full(edge.shape[1], states.ONE_ATM)])
    p = states.ONE_ATM
    out = {'species': np.array(sp), 'p_bk1': np.array(p), 'state_bk1': s_bk1,
           'state_bk2': s_bk2, 'p_bk2': p_bk2}
    for tag, ref in (('fast', ref_fast), ('strict', ref_strict)):
        out[f'rates_{tag}'] = ref.production_rates(s_bk1, p)
        c, v, rd = ref.transport(s_bk2, 1.0)
        out[f'cond_{tag}'], out[f'visc_{tag}'], out[f'rhoD_{tag}'] = c, v, rd
        rho, cp, rhocp = ref.thermo(s_bk1, p)
        out[f'rho_{tag}'], out[f'cp_{tag}'], out[f'rhocp_{tag}'] = rho, cp, rhocp
    np.savez_compressed(os.path.join(HERE, 'gri30_ref.npz'), **out)

    os.makedirs(os.path.join(HERE, 'ci_data'), exist_ok=True)
    for st in ('initial', 'ignition', 'final', 'ignition.highP'):
        src = os.path.join(REF, 'kinetix', 'ci_data', f'gri30.{st}.cantera')
        shutil.copy(src, os.path.join(HERE, 'ci_data', os.path.basename(src)))
    print('wrote', os.path.join(HERE, 'gri30_ref.npz'))


if __name__ == '__main__':
    main()
