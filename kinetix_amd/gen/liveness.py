"""Reaction order for the straight-line BK1 kernels, chosen for register liveness.

In the w-form kernels every species carries three FP64 values through the reaction stream:
its Gibbs term g_k (read by the exponents of the reactions it is a reactant of), w_k =
C_k exp(g_k) (read by the forward and reverse products) and its omega-dot accumulator. If
they are all computed up front and stored at the end, about 6N registers stay live for the
whole kernel (318 for GRI-3.0): past the 256 VGPRs of a wave, so the compiler spills to
AGPRs and the kernel runs at one wave per SIMD.

Emitting g_k and w_k at a species' first use and storing omega-dot_k after its last
reaction shortens every live range; the reaction order then decides the peak. This module
picks that order: greedy (next the reaction that opens the fewest species and closes the
most), then a seeded local search (move one reaction, keep the move if the peak and then
the summed profile do not grow). The register model counts 2 registers per FP64 value:

  c_k   from the start (the third-body sums read every C_k) to the first use of w_k;
  w_k   from its first use to its last use (reactant, or product of a reversible reaction);
  g_k   from the first use of w_k to its last use as a reactant;
  wd_k  from the first to the last reaction with a net stoichiometric coefficient for k.
"""
from __future__ import annotations

import random

import numpy as np


def use_masks(t):
    """(G, W, D): (R, N) boolean masks of the reactions reading g_k, w_k and writing wd_k."""
    nf, nr = np.asarray(t['nu_f']), np.asarray(t['nu_r'])
    rev = np.asarray(t['reversible']).astype(bool)[:, None]
    G = nf > 0
    W = G | (rev & (nr > 0))
    D = nr != nf
    return G, W, D


def _first_last(mask_ordered):
    """First and last row index of every column that has a True entry (-1 where none)."""
    R = mask_ordered.shape[0]
    any_ = mask_ordered.any(0)
    first = np.where(any_, mask_ordered.argmax(0), -1)
    last = np.where(any_, R - 1 - mask_ordered[::-1].argmax(0), -1)
    return first, last, any_


def profile(masks, order):
    """Modelled live FP64 registers before each reaction of `order`."""
    G, W, D = (m[order] for m in masks)
    R, N = G.shape
    delta = np.zeros(R + 1)
    fw, lw, aw = _first_last(W)
    fg, lg, ag = _first_last(G)
    fd, ld, ad = _first_last(D)
    np.add.at(delta, fw[aw], 2.0)
    np.add.at(delta, lw[aw] + 1, -2.0)
    ag = ag & aw
    np.add.at(delta, fw[ag], 2.0)
    np.add.at(delta, lg[ag] + 1, -2.0)
    np.add.at(delta, fd[ad], 2.0)
    np.add.at(delta, ld[ad] + 1, -2.0)
    delta[0] += 2.0 * N
    np.add.at(delta, np.where(aw, fw, 0), -2.0)
    return np.cumsum(delta)[:R]


def cost(masks, order):
    p = profile(masks, np.asarray(order))
    return float(p.max()), float(p.sum())


def greedy(masks):
    G, W, D = masks
    R, N = G.shape
    used = G | W | D
    remaining = used.sum(0).astype(int)
    started = np.zeros(N, bool)
    left = list(range(R))
    order = []
    while left:
        best, best_r = None, None
        for r in left:
            u = used[r]
            opens = int((u & ~started).sum())
            closes = int((u & (remaining == 1)).sum())
            key = (opens - closes, r)
            if best is None or key < best:
                best, best_r = key, r
        left.remove(best_r)
        order.append(best_r)
        started |= used[best_r]
        remaining -= used[best_r]
    return order


def order_reactions(t, iters=20000, seed=20261202):
    """Reaction order (list of reaction indices) with a low modelled register peak."""
    masks = use_masks(t)
    order = greedy(masks)
    R = len(order)
    cur = cost(masks, order)
    rng = random.Random(seed)
    for _ in range(iters):
        i, j = rng.randrange(R), rng.randrange(R)
        if i == j:
            continue
        trial = order[:]
        trial.insert(j, trial.pop(i))
        c = cost(masks, trial)
        if c <= cur:
            order, cur = trial, c
    return order
