"""Problem set-up with the reference's option API.

Mirrors ad_afqmc/mpi_jax.py ``_prep_afqmc`` (:39-140 integrals, trial orbitals and
options, :314-352 propagator / sampler selection) for the phaseless RHF / UHF path:
given the integrals of an FCIDUMP_chol file (pyscf_interface.py:578-598 writes them:
header [nelec, nmo, ms, nchol], energy_core, hcore, chol flat) and the trial orbitals
of mo_coeff.npz, it returns (ham_data, ham, prop, trial, wave_data, sampler,
observable, options) ready for driver.afqmc.

``read_fcidump_chol`` reads the HDF5 file (its own reader, no h5py needed) or an .npz
holding the same datasets; ``read_options`` reads an options.bin file.
"""

import numpy as np

from . import hamiltonian, hdf5_read, propagation, sampling, wavefunctions

OPTION_DEFAULTS = {
    # mpi_jax.py:62-90
    "dt": 0.01,
    "n_exp_terms": 6,
    "n_walkers": 50,
    "n_prop_steps": 50,
    "n_ene_blocks": 50,
    "n_sr_blocks": 1,
    "n_blocks": 50,
    "n_eql": 1,
    "ad_mode": None,
    "walker_type": "rhf",
    "trial": None,
    "ene0": 0.0,
    "n_batch": 1,
    # not in the reference: arithmetic of the Cholesky contractions ("fp64" or "bf16x3")
    "precision": "fp64",
}


def read_options(path):
    """The options dict of an options.bin file (mpi_jax.py:55-60 reads it with pickle.load;
    run_afqmc.py:31-32 writes it with pickle.dump).

    Decoded without the pickle module: binary pickle protocols 2-5 holding only dicts,
    lists, tuples, str, bytes, int, float, bool and None.  A file that needs anything
    else (a class, a global, a reduce call) raises ValueError; nothing in the file is
    executed.  A missing file gives {} (mpi_jax.py:57-61 falls back to defaults)."""
    try:
        with open(path, "rb") as f:
            data = f.read()
    except FileNotFoundError:
        return {}
    return _decode_plain_pickle(data)


def _decode_plain_pickle(data):
    import struct

    stack, marks, memo = [], [], {}
    pos, n = 0, len(data)

    def take(k):
        nonlocal pos
        if pos + k > n:
            raise ValueError("truncated options pickle")
        b = data[pos:pos + k]
        pos += k
        return b

    def pop_mark():
        m = marks.pop()
        items = stack[m:]
        del stack[m:]
        return items

    while True:
        op = take(1)[0]
        if op == 0x80:      # PROTO
            if take(1)[0] < 2:
                raise ValueError("options pickle: protocol < 2 is not supported")
        elif op == 0x95:    # FRAME
            take(8)
        elif op == 0x2E:    # STOP
            if len(stack) != 1:
                raise ValueError("malformed options pickle")
            return stack[0]
        elif op == 0x7D:    # EMPTY_DICT
            stack.append({})
        elif op == 0x5D:    # EMPTY_LIST
            stack.append([])
        elif op == 0x29:    # EMPTY_TUPLE
            stack.append(())
        elif op == 0x28:    # MARK
            marks.append(len(stack))
        elif op == 0x75:    # SETITEMS
            items = pop_mark()
            d = stack[-1]
            for i in range(0, len(items), 2):
                d[items[i]] = items[i + 1]
        elif op == 0x73:    # SETITEM
            v = stack.pop()
            k = stack.pop()
            stack[-1][k] = v
        elif op == 0x65:    # APPENDS
            items = pop_mark()
            stack[-1].extend(items)
        elif op == 0x61:    # APPEND
            v = stack.pop()
            stack[-1].append(v)
        elif op == 0x74:    # TUPLE
            stack.append(tuple(pop_mark()))
        elif op in (0x85, 0x86, 0x87):  # TUPLE1/2/3
            k = op - 0x84
            t = tuple(stack[-k:])
            del stack[-k:]
            stack.append(t)
        elif op == 0x8C:    # SHORT_BINUNICODE
            stack.append(take(take(1)[0]).decode("utf-8"))
        elif op == 0x58:    # BINUNICODE
            stack.append(take(struct.unpack("<I", take(4))[0]).decode("utf-8"))
        elif op == 0x8D:    # BINUNICODE8
            stack.append(take(struct.unpack("<Q", take(8))[0]).decode("utf-8"))
        elif op == 0x43:    # SHORT_BINBYTES
            stack.append(bytes(take(take(1)[0])))
        elif op == 0x42:    # BINBYTES
            stack.append(bytes(take(struct.unpack("<I", take(4))[0])))
        elif op == 0x4B:    # BININT1
            stack.append(take(1)[0])
        elif op == 0x4D:    # BININT2
            stack.append(struct.unpack("<H", take(2))[0])
        elif op == 0x4A:    # BININT
            stack.append(struct.unpack("<i", take(4))[0])
        elif op == 0x8A:    # LONG1
            stack.append(int.from_bytes(take(take(1)[0]), "little", signed=True))
        elif op == 0x47:    # BINFLOAT
            stack.append(struct.unpack(">d", take(8))[0])
        elif op == 0x88:    # NEWTRUE
            stack.append(True)
        elif op == 0x89:    # NEWFALSE
            stack.append(False)
        elif op == 0x4E:    # NONE
            stack.append(None)
        elif op == 0x94:    # MEMOIZE
            memo[len(memo)] = stack[-1]
        elif op == 0x71:    # BINPUT
            memo[take(1)[0]] = stack[-1]
        elif op == 0x72:    # LONG_BINPUT
            memo[struct.unpack("<I", take(4))[0]] = stack[-1]
        elif op == 0x68:    # BINGET
            stack.append(memo[take(1)[0]])
        elif op == 0x6A:    # LONG_BINGET
            stack.append(memo[struct.unpack("<I", take(4))[0]])
        else:
            raise ValueError(f"options pickle: opcode 0x{op:02x} would construct an object "
                             "beyond plain values; refusing to decode it")


def read_fcidump_chol(path):
    """Datasets of an FCIDUMP_chol file as numpy arrays (mpi_jax.py:39-51): the HDF5 file
    through the library's own reader (hdf5_read; h5py, when installed, for files beyond
    the layouts that reader handles), or an .npz holding the same datasets."""
    if str(path).endswith(".npz"):
        with np.load(path) as z:
            return {k: np.array(z[k]) for k in z.files}
    try:
        return hdf5_read.read_datasets(path)
    except ValueError:
        try:
            import h5py
        except ImportError:
            raise
    with h5py.File(path, "r") as f:
        return {k: np.array(f[k]) for k in f.keys()}


def prep_afqmc_files(options=None, option_file="options.bin", mo_file="mo_coeff.npz",
                     chol_file="FCIDUMP_chol"):
    """mpi_jax.py:36-60 with its file arguments: the integrals of chol_file (HDF5 or .npz),
    the trial orbitals mo_file["mo_coeff"] and, when options is None, the options of
    option_file (plain values only, see read_options); then prep_afqmc."""
    if options is None:
        options = read_options(option_file)
    with np.load(mo_file) as z:
        mo_coeff = np.array(z["mo_coeff"])
    return prep_afqmc(read_fcidump_chol(chol_file), mo_coeff, options)


def prep_afqmc(integrals, mo_coeff, options=None):
    """mpi_jax.py:39-140 and :314-352 for trial "rhf" / "uhf".

    integrals: dict with header [nelec, nmo, ms, nchol], energy_core, hcore (nmo*nmo),
    chol (nchol*nmo*nmo); mo_coeff: (2, nmo, nmo) trial orbitals (mo_coeff.npz)."""
    options = dict(OPTION_DEFAULTS, **(options or {}))
    if options["ad_mode"] is not None:
        raise NotImplementedError("only ad_mode=None is supported")
    nelec, nmo, ms, nchol = (int(x) for x in np.asarray(integrals["header"]).ravel())
    nelec_sp = ((nelec + abs(ms)) // 2, (nelec - abs(ms)) // 2)
    h0 = float(np.asarray(integrals["energy_core"]))
    h1 = np.asarray(integrals["hcore"], dtype=np.float64).reshape(nmo, nmo)
    chol = np.asarray(integrals["chol"], dtype=np.float64).reshape(nchol, nmo * nmo)
    ham_data = {"h0": h0, "h1": np.array([h1, h1]), "chol": chol, "ene0": options["ene0"]}
    mo_coeff = np.asarray(mo_coeff, dtype=np.float64)
    wave_data = {"rdm1": np.array([
        mo_coeff[0][:, :nelec_sp[0]] @ mo_coeff[0][:, :nelec_sp[0]].T,
        mo_coeff[1][:, :nelec_sp[1]] @ mo_coeff[1][:, :nelec_sp[1]].T])}
    trial_kind = options["trial"] or options["walker_type"]
    if trial_kind == "rhf":
        trial = wavefunctions.rhf(nmo, nelec_sp, n_batch=options["n_batch"])
        wave_data["mo_coeff"] = mo_coeff[0][:, :nelec_sp[0]]
    elif trial_kind == "uhf":
        trial = wavefunctions.uhf(nmo, nelec_sp, n_batch=options["n_batch"])
        wave_data["mo_coeff"] = [mo_coeff[0][:, :nelec_sp[0]], mo_coeff[1][:, :nelec_sp[1]]]
    else:
        raise NotImplementedError(f"trial {trial_kind!r}: only rhf / uhf are supported")
    ham = hamiltonian.hamiltonian(nmo)
    if options["walker_type"] == "rhf":
        prop = propagation.propagator_restricted(
            options["dt"], options["n_walkers"], options["n_exp_terms"], options["n_batch"],
            precision=options["precision"])
    elif options["walker_type"] == "uhf":
        prop = propagation.propagator_unrestricted(
            options["dt"], options["n_walkers"], options["n_exp_terms"], options["n_batch"],
            precision=options["precision"])
    else:
        raise ValueError(f"walker_type {options['walker_type']!r}")
    smp = sampling.sampler(options["n_prop_steps"], options["n_ene_blocks"],
                           options["n_sr_blocks"], options["n_blocks"])
    observable = None
    return ham_data, ham, prop, trial, wave_data, smp, observable, options
