"""Batched fragments: the grouped step against each fragment's own single-context run,
and the independence of a fragment's result from the rest of the batch.

Tolerances: float64 groups agree with the single-context sampler to 1e-9 relative in
block energies (the single path splits the force-bias GEMM over K, the grouped one does
not); mixed-precision groups to 1e-5.  Alone vs in a mix: bit-identical."""

import numpy as np
import pytest
import torch

from cs_afqmc_amd import batch, problems, rng, sampling

pytestmark = pytest.mark.gpu

SPECS = [
    # walker_type, norb, nelec, nchol, W, seed
    ("rhf", 20, (4, 4), 80, 32, 1),
    ("uhf", 36, (7, 5), 90, 16, 2),
    ("rhf", 57, (11, 11), 228, 24, 3),
    ("uhf", 140, (30, 28), 150, 8, 4),  # streaming Taylor kernels (N > 128)
    ("rhf", 100, (20, 20), 400, 16, 5),
]


def make_problem(spec):
    wt, norb, nelec, nchol, W, seed = spec
    cs = 0.5 / np.sqrt(norb)
    if wt == "rhf":
        hd, wd, ne = problems.synthetic_rhf(norb, nelec[0], nchol, seed=seed, chol_std=cs)
    else:
        hd, wd, ne = problems.synthetic_uhf(norb, nelec[0], nelec[1], nchol, seed=seed,
                                            chol_std=cs)
    return hd, wd, ne


def fragment(spec, precision, n_prop_steps):
    wt, norb, nelec, nchol, W, seed = spec
    hd, wd, ne = make_problem(spec)
    return batch.Fragment(hd, wd, wt, ne, W, dt=0.005, seed=seed, precision=precision,
                          n_prop_steps=n_prop_steps, tag=spec)


def walkers_np(pd):
    w = pd["walkers"]
    return [x.cpu().numpy() for x in w] if isinstance(w, list) else [w.cpu().numpy()]


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("bf16x3", 1e-5)])
def test_group_matches_single_context_sampler(gpu, precision, tol):
    steps, blocks = 10, 2
    frags = [fragment(s, precision, steps) for s in SPECS]
    b = batch.FragmentBatch(frags)
    for _ in range(blocks):
        b.run_block()
    for spec, f in zip(SPECS, frags):
        print(f"{precision} fragment {spec}", flush=True)
        ref = fragment(spec, precision, steps)
        # the fragment driver: one-block sampler calls, e_estimate <- 0.9 e + 0.1 E after
        # each (lno_afqmc/ccsd_pt2/run_afqmc_nompi.py:83-89)
        smp = sampling.sampler(n_prop_steps=steps, n_ene_blocks=1, n_sr_blocks=1)
        pd = ref.prop_data
        be_ref = []
        for _ in range(blocks):
            e_ref, pd = smp.propagate_phaseless(None, ref.ham_data, ref.prop, pd, ref.trial,
                                                ref.wave_data)
            be_ref.append(pd["block_energies"].cpu().numpy()[0])
            pd["e_estimate"] = 0.9 * pd["e_estimate"] + 0.1 * e_ref
        be_ref = np.array(be_ref)
        be, _ = f.block_energies()
        assert np.all(np.isfinite(be))
        scale = np.max(np.abs(be_ref))
        assert np.max(np.abs(be - be_ref)) <= tol * scale
        for a, r in zip(walkers_np(f.prop_data), walkers_np(pd)):
            assert np.max(np.abs(a - r)) <= 100 * tol * np.max(np.abs(r))


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_fragment_alone_and_in_mix_bit_identical(gpu, precision):
    steps = 8
    target = SPECS[2]
    alone = fragment(target, precision, steps)
    b1 = batch.FragmentBatch([alone])
    b1.run_block()
    mix = [fragment(s, precision, steps) for s in SPECS]
    b2 = batch.FragmentBatch(mix, tune=True)
    b2.run_block()
    other = mix[2]
    e1, w1 = alone.block_energies()
    e2, w2 = other.block_energies()
    assert np.array_equal(e1, e2) and np.array_equal(w1, w2)
    for a, r in zip(walkers_np(alone.prop_data), walkers_np(other.prop_data)):
        assert np.array_equal(a, r)


def test_batch_membership_changes(gpu):
    """Fragments leave and join between blocks; the survivors keep their trajectory."""
    steps = 6
    frags = [fragment(s, "bf16x3", steps) for s in SPECS[:3]]
    ref = fragment(SPECS[1], "bf16x3", steps)
    b = batch.FragmentBatch(frags)
    rb = batch.FragmentBatch([ref])
    b.run_block()
    rb.run_block()
    b.remove([frags[0]])
    b.add(fragment(SPECS[4], "bf16x3", steps))
    b.run_block()
    rb.run_block()
    assert np.array_equal(frags[1].block_energies()[0], ref.block_energies()[0])
    assert len(b.group) == 3


def test_graph_replay_matches_eager(gpu):
    """Blocks replayed from a captured graph equal eager blocks bit for bit; the replays
    are interleaved with the fragments' own eager block-end work and synchronisations."""
    steps, blocks = 6, 4
    eager = [fragment(s, "bf16x3", steps) for s in SPECS[:4]]
    graph = [fragment(s, "bf16x3", steps) for s in SPECS[:4]]
    side = torch.cuda.Stream()
    be = batch.FragmentBatch(eager, stream=side)
    bg = batch.FragmentBatch(graph, use_graph=True, stream=side)
    for b in range(blocks):
        be.run_block()
        bg.run_block()
        torch.cuda.synchronize()
    assert bg._graph is not None  # blocks 2.. replayed the graph
    for fe, fg in zip(eager, graph):
        assert np.array_equal(fe.block_energies()[0], fg.block_energies()[0])
        for a, r in zip(walkers_np(fe.prop_data), walkers_np(fg.prop_data)):
            assert np.array_equal(a, r)


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_grouped_block_ends_match_each_context(gpu, precision):
    """afqmc_group_orthonormalize / _overlap / _energy against each member context's own
    afqmc_orthonormalize / _overlap / _energy on the same walkers: QR and overlaps bit for
    bit, energies to 1e-12 relative (the grouped force-bias part c_g is one float64
    product per spin, the context's splits it over K)."""
    frags = [fragment(s, precision, 4) for s in SPECS]
    b = batch.FragmentBatch(frags)
    b.run_block()  # walkers away from the trial
    torch.cuda.synchronize()
    copies = [[w.clone() for w in (f.prop_data["walkers"]
                                    if isinstance(f.prop_data["walkers"], list)
                                    else [f.prop_data["walkers"]])] for f in frags]
    with torch.cuda.stream(b.stream):
        b.group.orthonormalize()
        b.group.overlap()
        en = [torch.empty_like(f.energies) for f in frags]
        b.group.energy(en)
    torch.cuda.synchronize()
    for f, ws, e in zip(frags, copies, en):
        ref = ws if len(ws) > 1 else ws[0]
        f.dev.orthonormalize(ref)
        ov = f.dev.overlap(ref)
        er = f.dev.energy(ref)
        torch.cuda.synchronize()
        for a, r in zip(walkers_np(f.prop_data), [w.cpu().numpy() for w in ws]):
            assert np.array_equal(a, r)
        assert torch.equal(f.prop_data["overlaps"], ov)
        e_np, r_np = e.cpu().numpy(), er.cpu().numpy()
        assert np.all(np.isfinite(e_np))
        assert np.max(np.abs(e_np - r_np)) <= 1e-12 * np.max(np.abs(r_np))


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_tuned_group_energy_matches_each_context(gpu, precision):
    """afqmc_group_tune times both tile edges of every grouped GEMM, the energy's float64
    c_g product included, on the bound buffers.  Afterwards the grouped energy (with the
    chosen c_g tile) still matches each context's own afqmc_energy to 1e-12 relative, and
    the grouped energy computed before tuning bit for bit (the k order of a GEMM output
    does not depend on the tile edge)."""
    frags = [fragment(s, precision, 4) for s in SPECS]
    b = batch.FragmentBatch(frags)
    b.run_block()
    torch.cuda.synchronize()
    with torch.cuda.stream(b.stream):
        before = [torch.empty_like(f.energies) for f in frags]
        b.group.energy(before)
    torch.cuda.synchronize()
    tuned = b.tune()
    assert set(tuned) == {"fb", "vhs", "cg"}
    assert all(t > 0 for pair in tuned.values() for t in pair)
    with torch.cuda.stream(b.stream):
        after = [torch.empty_like(f.energies) for f in frags]
        b.group.energy(after)
    torch.cuda.synchronize()
    for f, e0, e in zip(frags, before, after):
        assert torch.equal(e0, e)
        er = f.dev.energy(f.prop_data["walkers"])
        torch.cuda.synchronize()
        e_np, r_np = e.cpu().numpy(), er.cpu().numpy()
        assert np.all(np.isfinite(e_np))
        assert np.max(np.abs(e_np - r_np)) <= 1e-12 * np.max(np.abs(r_np))


def test_grouped_fields_draw_matches_each_context(gpu):
    """afqmc_group_fields_normal_key against each member's afqmc_fields_normal_key from a
    copy of its key: the fields blocks bit for bit, the keys advanced alike."""
    frags = [fragment(s, "bf16x3", 3) for s in SPECS]
    b = batch.FragmentBatch(frags)
    keys0 = [f.key.clone() for f in frags]
    with torch.cuda.stream(b.stream):
        b.group.fields_normal_key([f.key for f in frags])
    torch.cuda.synchronize()
    for f, k in zip(frags, keys0):
        ref = torch.empty_like(f.fields)
        f.dev.fields_normal_key(k, ref)
        torch.cuda.synchronize()
        assert torch.equal(f.fields, ref)
        assert torch.equal(f.key, k)
