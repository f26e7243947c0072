"""Generator CLI: ``python -m kinetix_amd.gen --mechanism M.yaml --output DIR [options]``.

Mirrors the reference generator's entry point (kinetix/__main__.py:22-38,
kinetix/utils/general_utils.py:20-87): reads a Cantera-YAML mechanism and writes the
mechanism tables (tables.npz, mech.json) and the gfx950 HIP kernels (kernels.hip) for it.
"""
import argparse
import os
import sys
import time

from . import mechanism, tables, transport_fit


def main(argv=None):
    ap = argparse.ArgumentParser(description='Generate MI355X (gfx950) kernels for a mechanism')
    ap.add_argument('--mechanism', required=True, help='path to the Cantera YAML mechanism')
    ap.add_argument('--output', required=True, help='output directory')
    ap.add_argument('--name', default=None, help='mechanism name (default: YAML stem)')
    ap.add_argument('--fit-rcpdiffcoeffs', action='store_true',
                    help='fit T^1.5/(p D_jk) instead of p D_jk/T^1.5 (no divides in BK2)')
    ap.add_argument('--no-transport', action='store_true', help='skip transport fits/kernels')
    ap.add_argument('--tables-only', action='store_true', help='write tables, no kernels')
    ap.add_argument('--rolled', action='store_true',
                    help='loop forms of the FP64/mixed kernels (default above 96 species)')
    ap.add_argument('--bk2-loop', action='store_true',
                    help='with the loop forms: transport one state per lane instead of state tiles')
    args = ap.parse_args(argv)

    t0 = time.time()
    mech = mechanism.load(args.mechanism, args.name)
    fits = None
    if not args.no_transport:
        fits = transport_fit.fit(mech, reciprocal_diffusivity=args.fit_rcpdiffcoeffs)
    tab = tables.build(mech, fits)
    tables.save(args.output, mech, tab)
    from . import counts
    counts.write(args.output, tab)
    if not args.tables_only:
        from . import emit_hip
        opts = {}
        if args.rolled:
            opts['rolled'] = True
        if args.bk2_loop:
            opts['bk2_loop'] = True
        emit_hip.write(args.output, mech, tab, opts or None)
    print(f'{mech.name}: {mech.n_species} species ({mech.n_active} active), '
          f'{mech.n_reactions} reactions -> {args.output} [{time.time() - t0:.1f}s]')
    return 0


if __name__ == '__main__':
    sys.exit(main())
