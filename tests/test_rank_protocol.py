"""A host model of the radix downsweep's in-wave ranking protocol
(endgraph_amd/csrc/radix.hpp, rx_rank): every lane ORs its bit into its
digit's entry of a per-wave LDS table of lane masks, reads the entry back,
clears it, then reads the wave's running count of the digit, and the highest
peer writes the count back.  Lanes whose digit is the first lane's take
their peers from a ballot instead and stay off the table (a wave with no
other digit skips the table); lead_ok=False models the table alone.  The table and the counts belong to one wave,
so other waves never touch them; what can go wrong is the order in which one
wave's LDS accesses complete.

Two memory semantics are modelled:
  in-order    the LDS completes a wave's instructions in issue order (what
              gfx9-family hardware does);
  waits-only  accesses issued after the last s_waitcnt lgkmcnt(0) may
              complete in any order (only the waits order them: what the
              source may assume without relying on the hardware).
Each instruction acts on all active lanes at once.  The protocol is checked
against the stable ranks under every completion order the semantics allow
(enumerated per window between waits), and the waits are dropped one at a
time to show which of them the result depends on.  Not a GPU test: the
device code is exercised by the build's parity tests."""
import itertools

import numpy as np
import pytest

WAVE = 64

# One record of the loop as issued, "W" for s_waitcnt lgkmcnt(0): OR, wait,
# read the mask, wait, clear, read the count, wait, write the count (all
# three waits are written in the source).
SHIPPED = ["or", "W", "read", "W", "clear", "rcount", "W", "wcount"]


def expected_ranks(dig, valid):
    """Stable rank of each valid record among the wave's records of its digit:
    records in order k, then lane."""
    seen = {}
    out = np.zeros(dig.shape, np.int64)
    for k in range(dig.shape[0]):
        for lane in range(WAVE):
            if valid[k, lane]:
                d = int(dig[k, lane])
                out[k, lane] = seen.get(d, 0)
                seen[d] = seen.get(d, 0) + 1
    return out


def windows(seq, n_records):
    """Issue stream of all records split into windows at each wait."""
    ops = [(k, op) for k in range(n_records) for op in seq]
    out, cur = [], []
    for k, op in ops:
        if op == "W":
            out.append(cur)
            cur = []
        else:
            cur.append((k, op))
    out.append(cur)
    return [w for w in out if w]


def run(dig, valid, seq, order_of, lead_ok=True):
    """Executes the stream; order_of(window) gives the completion order of
    a window's accesses.  Returns the ranks per (record, lane)."""
    tab, count = {}, {}
    peers = np.zeros(dig.shape, np.uint64)
    prior = np.zeros(dig.shape, np.int64)
    rank = np.zeros(dig.shape, np.int64)
    lead = np.zeros(dig.shape, bool)
    for k in range(dig.shape[0]):
        # the ballot of the lanes sharing lane 0's digit (lane 0 is the first
        # active lane: every lane of the wave runs the loop)
        lead[k] = valid[k] & (dig[k] == dig[k, 0]) & lead_ok
        m0 = sum(1 << l for l in range(WAVE) if lead[k, l])
        for l in range(WAVE):
            if lead[k, l]:
                peers[k, l] = m0
    for win in windows(seq, dig.shape[0]):
        for k, op in order_of(win):
            ds = dig[k]
            lanes = [l for l in range(WAVE) if valid[k, l]]
            tabbed = [l for l in lanes if not lead[k, l]]
            if op == "or":
                for l in tabbed:
                    tab[ds[l]] = tab.get(ds[l], 0) | (1 << l)
            elif op == "read":
                for l in tabbed:
                    peers[k, l] = tab.get(ds[l], 0)
            elif op == "clear":
                for l in tabbed:
                    tab[ds[l]] = 0
            elif op == "rcount":
                for l in range(WAVE):
                    prior[k, l] = count.get(ds[l], 0)
            elif op == "wcount":
                for l in lanes:
                    p = int(peers[k, l])
                    below = bin(p & ((1 << l) - 1)).count("1")
                    total = bin(p).count("1")
                    if below + 1 == total:
                        count[ds[l]] = int(prior[k, l]) + total
    for k in range(dig.shape[0]):
        for l in range(WAVE):
            p = int(peers[k, l])
            rank[k, l] = prior[k, l] + bin(p & ((1 << l) - 1)).count("1")
    return rank


def respects_data(order):
    """A count write needs its record's peers and prior count: the compiler
    waits for those reads whatever the source says."""
    pos = {a: i for i, a in enumerate(order)}
    for (k, op), i in pos.items():
        if op == "wcount":
            for dep in ("read", "rcount"):
                if (k, dep) in pos and pos[(k, dep)] > i:
                    return False
    return True


def all_orders(dig, valid, seq, limit=2000, lead_ok=True):
    """Ranks under every completion order the waits-only semantics allow
    (the product of the windows' permutations, capped)."""
    wins = windows(seq, dig.shape[0])
    perms = [[c for c in itertools.permutations(w) if respects_data(c)] for w in wins]
    results = []
    for n, choice in enumerate(itertools.product(*perms)):
        if n >= limit:
            break
        table = {tuple(w): list(c) for w, c in zip(wins, choice)}
        results.append(run(dig, valid, seq, lambda win: table[tuple(win)], lead_ok))
    return results


def cases():
    rng = np.random.default_rng(5)
    out = []
    # few records (the enumeration is exponential in them), digits that
    # repeat across consecutive records of a lane, a ragged tail
    for nb in (1, 2, 4):
        dig = rng.integers(0, nb, (3, WAVE))
        valid = np.ones(dig.shape, bool)
        out.append((dig, valid))
    dig = rng.integers(0, 3, (3, WAVE))
    valid = np.ones(dig.shape, bool)
    valid[2, 40:] = False
    out.append((dig, valid))
    # a run of lane 0's digit over most of the wave, other digits around it
    dig = rng.integers(1, 4, (3, WAVE))
    dig[:, :50] = 0
    dig[1, 10] = 2
    valid = np.ones(dig.shape, bool)
    out.append((dig, valid))
    return out


@pytest.mark.parametrize("lead_ok", [True, False])
@pytest.mark.parametrize("case", range(5))
def test_shipped_sequence_in_order(case, lead_ok):
    dig, valid = cases()[case]
    got = run(dig, valid, SHIPPED, lambda w: w, lead_ok)
    exp = expected_ranks(dig, valid)
    assert np.array_equal(got[valid], exp[valid])


@pytest.mark.parametrize("lead_ok", [True, False])
@pytest.mark.parametrize("case", range(5))
def test_shipped_sequence_under_every_waits_only_order(case, lead_ok):
    dig, valid = cases()[case]
    exp = expected_ranks(dig, valid)
    for got in all_orders(dig, valid, SHIPPED, lead_ok=lead_ok):
        assert np.array_equal(got[valid], exp[valid])


def drop_wait(seq, i):
    idx = [j for j, op in enumerate(seq) if op == "W"][i]
    return seq[:idx] + seq[idx + 1:]


@pytest.mark.parametrize("which,breaks", [(0, "read before the ORs"), (1, "clear before the read"),
                                          (2, "clear lands after the next record's OR")])
def test_each_wait_is_needed_without_in_order_lds(which, breaks):
    """Dropping any one wait lets some waits-only completion order give wrong
    ranks; the third is the subtle one: the clear of record k and the OR of
    record k+1 on the same digit then share a window."""
    # every lane but lane 0 shares one digit: they all collide in the table
    dig = np.ones((3, WAVE), np.int64)
    dig[:, 0] = 0
    valid = np.ones(dig.shape, bool)
    exp = expected_ranks(dig, valid)
    seq = drop_wait(SHIPPED, which)
    assert any(not np.array_equal(g, exp) for g in all_orders(dig, valid, seq)), breaks
    # in issue order the same stream is still right: the hardware's LDS
    # ordering alone would carry it
    assert np.array_equal(run(dig, valid, seq, lambda w: w), exp)
