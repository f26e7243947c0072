"""kx_init_image (the host-application hook: mechanism text and code object from memory)
against kx_init from the directory, through the C ABI: bit-identical outputs."""
import os

import numpy as np
import pytest
import torch

from conftest import ROOT
from kinetix_amd import states
from kinetix_amd.api import Kinetix

pytestmark = pytest.mark.gpu


def test_image_init_matches_directory_init():
    d = os.path.join(ROOT, 'kinetix_amd', 'mechs', 'gri30')
    image = (open(os.path.join(d, 'mech.txt'), 'rb').read(),
             open(os.path.join(d, 'kernels.hsaco'), 'rb').read())
    a, b = Kinetix('gri30'), Kinetix(image=image)
    try:
        assert b.speciesNames() == a.speciesNames()
        sp, W = a.speciesNames(), a.molar_masses
        for kx in (a, b):
            kx.build(101325.0, 1.0, states.reference_mass_fractions(sp, W))
        n, N = 4096, a.nSpecies()
        s, _ = states.sample(n, sp, W, seed=states.BASE_SEED + 9)
        ds = torch.from_numpy(np.ascontiguousarray(s)).cuda().reshape(-1)
        outs = []
        for kx in (a, b):
            r = torch.full_like(ds, np.nan)
            kx.productionRates(n, n, n, 1.0, ds, r)
            v, c = (torch.full((n,), np.nan, dtype=torch.float64, device='cuda') for _ in range(2))
            rd = torch.full((N * n,), np.nan, dtype=torch.float64, device='cuda')
            kx.mixtureAvgTransportProps(n, n, n, 1.0, ds, v, c, rd)
            torch.cuda.synchronize()
            outs.append((r, v, c, rd))
        for x, y in zip(*outs):
            assert torch.equal(x, y)
    finally:
        a.close()
        b.close()
