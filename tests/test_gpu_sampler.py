"""The device-resident sampler (sampling.propagate_phaseless) against the oracle's
propagate_phaseless (sampling.py:319-341 of the reference) driven by the same
auxiliary fields and SR draws; the device field generator against Philox4x32-10."""

import numpy as np
import pytest
import torch

from cs_afqmc_amd import hamiltonian, problems, propagation, rng, sampling, wavefunctions
from oracle import afqmc_oracle as o

pytestmark = pytest.mark.gpu


def rel(a, b):
    if isinstance(b, (list, tuple)):
        return max(rel(x, y) for x, y in zip(a, b))
    a, b = np.asarray(a), np.asarray(b)
    s = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / (s if s > 0 else 1.0)


def to_np(x):
    if isinstance(x, (list, tuple)):
        return [to_np(y) for y in x]
    return x.detach().cpu().numpy()


@pytest.mark.parametrize("walker_type", ["rhf", "uhf"])
def test_sampler_matches_oracle(gpu, walker_type):
    norb, nchol, W, dt = 14, 30, 16, 0.01
    print(f"sampler case {walker_type} N={norb} G={nchol} W={W}", flush=True)
    cs = 0.5 / np.sqrt(norb)
    if walker_type == "rhf":
        hd, wd, nelec = problems.synthetic_rhf(norb, 4, nchol, seed=21, chol_std=cs)
        trial = wavefunctions.rhf(norb, nelec)
        prop = propagation.propagator_restricted(dt=dt, n_walkers=W)
    else:
        hd, wd, nelec = problems.synthetic_uhf(norb, 4, 3, nchol, seed=21, chol_std=cs)
        trial = wavefunctions.uhf(norb, nelec)
        prop = propagation.propagator_unrestricted(dt=dt, n_walkers=W)
    hd["h0"] = 0.2
    smp = sampling.sampler(n_prop_steps=5, n_ene_blocks=2, n_sr_blocks=2)
    ham = hamiltonian.hamiltonian(norb)
    wave = dict(wd)
    hdat = ham.build_measurement_intermediates(dict(hd), trial, wave)
    hdat = ham.build_propagation_intermediates(hdat, prop, trial, wave)
    pd = prop.init_prop_data(trial, wave, hdat)
    pd["key"] = rng.PRNGKey(123)
    dev = hdat["_afqmc_device"]

    # the same fields and SR draws, regenerated from the key sequence the sampler uses
    key = pd["key"]
    fields = np.zeros((2, 2, 5, W, nchol))
    zetas = []
    for s in range(2):
        for b in range(2):
            key, sub = rng.split(key)
            fields[s, b] = to_np(dev.fields_normal(sub[0], sub[1], (5, W, nchol)))
        key, sub = rng.split(key)
        zetas.append(rng.uniform(sub))

    oh = o.build_measurement_intermediates(walker_type, dict(hd), wd, norb)
    oh = o.build_propagation_intermediates(walker_type, oh, wd["rdm1"], norb, nelec, dt)
    opd = o.init_prop_data(walker_type, walker_type, wd, oh, W, nelec)
    assert abs(float(pd["e_estimate"]) - opd["e_estimate"]) < 1e-10 * abs(opd["e_estimate"])

    e_gpu, pd = smp.propagate_phaseless(ham, hdat, prop, pd, trial, wave)
    e_ref, opd, be, bw = o.propagate_phaseless(walker_type, walker_type, oh, opd, wd, fields,
                                               zetas, dt, 6, norb, 2, 2)
    assert rel(to_np(pd["block_energies"]), be) < 1e-10
    assert rel(to_np(pd["block_weights"]), bw) < 1e-10
    assert abs(e_gpu - e_ref) < 1e-10 * abs(e_ref)
    assert rel(to_np(pd["walkers"]), opd["walkers"]) < 1e-10
    assert rel(to_np(pd["weights"]), opd["weights"]) < 1e-10
    assert abs(float(pd["pop_control_ene_shift"]) - opd["pop_control_ene_shift"]) < 1e-9


def test_device_normals_match_philox(gpu):
    """afqmc_fields_normal = Box-Muller on Philox4x32-10 words (host restatement)."""
    norb, nchol = 6, 5
    hd, wd, nelec = problems.synthetic_rhf(norb, 2, nchol, seed=1)
    trial = wavefunctions.rhf(norb, nelec)
    prop = propagation.propagator_restricted(dt=0.01, n_walkers=4)
    ham = hamiltonian.hamiltonian(norb)
    hdat = ham.build_measurement_intermediates(dict(hd), trial, dict(wd))
    hdat = ham.build_propagation_intermediates(hdat, prop, trial, dict(wd))
    dev = hdat["_afqmc_device"]
    seed, counter, count = 0x1234567890ABCDEF, 77, 301
    got = to_np(dev.fields_normal(seed, counter, (count,)))
    ref = np.empty(count)
    for pair in range((count + 1) // 2):
        r = rng.philox4x32_10((pair & 0xFFFFFFFF, pair >> 32, counter & 0xFFFFFFFF,
                               counter >> 32), (seed & 0xFFFFFFFF, seed >> 32))
        a = (r[1] << 32) | r[0]
        b = (r[3] << 32) | r[2]
        u1 = ((a >> 11) + 1) * 2.0**-53
        u2 = (b >> 11) * 2.0**-53
        rad = np.sqrt(-2.0 * np.log(u1))
        ref[2 * pair] = rad * np.cos(2 * np.pi * u2)
        if 2 * pair + 1 < count:
            ref[2 * pair + 1] = rad * np.sin(2 * np.pi * u2)
    assert np.max(np.abs(got - ref)) < 1e-13
    assert abs(np.mean(got)) < 0.2 and abs(np.std(got) - 1) < 0.15
