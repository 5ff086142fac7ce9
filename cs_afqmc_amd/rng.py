"""Counter-based random numbers for the sampler.

Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
1, 2, 3", SC'11).  The device kernel behind afqmc_fields_normal uses the same
bijection; the host side here draws the scalar uniform zeta of each stochastic
reconfiguration (the role of jax.random.uniform in propagation.py:234-241).

A key is the pair (seed, counter); ``split`` returns the key for the next draw and a
subkey, mirroring ``random.split`` usage in the reference.
"""

_M0, _M1 = 0xD2511F53, 0xCD9E8D57
_W0, _W1 = 0x9E3779B9, 0xBB67AE85
_MASK = 0xFFFFFFFF


def philox4x32_10(counter, key):
    """counter: 4 uint32, key: 2 uint32 -> 4 uint32."""
    c0, c1, c2, c3 = (int(x) & _MASK for x in counter)
    k0, k1 = (int(x) & _MASK for x in key)
    for _ in range(10):
        p0 = _M0 * c0
        p1 = _M1 * c2
        hi0, lo0 = p0 >> 32, p0 & _MASK
        hi1, lo1 = p1 >> 32, p1 & _MASK
        c0, c1, c2, c3 = (hi1 ^ c1 ^ k0) & _MASK, lo1, (hi0 ^ c3 ^ k1) & _MASK, lo0
        k0 = (k0 + _W0) & _MASK
        k1 = (k1 + _W1) & _MASK
    return c0, c1, c2, c3


def PRNGKey(seed):
    return (int(seed) & (2**64 - 1), 0)


def split(key):
    seed, counter = key
    return (seed, counter + 1), (seed, counter)


def _to_i64(u):
    u &= 2**64 - 1
    return u - 2**64 if u >= 2**63 else u


def device_key(key, device, out=None):
    """The key as a device int64[2] tensor {seed, counter} (uint64 bit patterns), the
    form afqmc_fields_normal_key / afqmc_sr_key read and advance on the device."""
    import torch

    host = torch.tensor([_to_i64(key[0]), _to_i64(key[1])], dtype=torch.int64)
    if out is None:
        return host.to(device)
    out.copy_(host)
    return out


def advance(key, n):
    """The key after n device draws (each fields block and each SR advances the
    counter by one, as one split does on the host)."""
    return (key[0], key[1] + int(n))


def uniform(key):
    """One float64 uniform in [0, 1) from the 64 high-quality bits of a Philox draw;
    stream 0xFFFFFFFF keeps it apart from the field draws of the same counter."""
    seed, counter = key
    r = philox4x32_10((0, 0xFFFFFFFF, counter & _MASK, (counter >> 32) & _MASK),
                      (seed & _MASK, (seed >> 32) & _MASK))
    a = (r[1] << 32) | r[0]
    return (a >> 11) * 2.0**-53
