"""The package's HDF5 reader and writer (mma_amd/h5lite.py), which stands in for h5py: write ->
read round trips of every supported kind (float64/32, int32/64, fixed-length strings, scalars,
attributes, nested groups, a group with more children than one symbol-table node holds), and,
where the HDF5 library's command-line tools are installed, the same files read by the HDF5
library itself (h5dump: structure, exact binary data, attributes) and a copy rewritten by the
library (h5repack) read back by h5lite.  Also the archive-facing functions that need no device:
the CTDSE inputs of an archive (ReadInputs, tools_hdf5.c:65-148) and the save/load of the
ctypes mirror's inputs and outputs (PythonCTDSE/structures.py:133-175,301-355,454-546)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from mma_amd import farm, h5lite

TOOLS = "/opt/conda/bin"
H5DUMP = shutil.which("h5dump") or (os.path.join(TOOLS, "h5dump") if os.path.exists(os.path.join(TOOLS, "h5dump"))
                                    else None)
H5REPACK = shutil.which("h5repack") or (os.path.join(TOOLS, "h5repack")
                                        if os.path.exists(os.path.join(TOOLS, "h5repack")) else None)


def _sample():
    rng = np.random.default_rng(3)
    d = {"/CTDSE/outputs/SourceTerm": rng.standard_normal((3, 4, 5)),
         "/CTDSE/outputs/FSourceTerm": rng.standard_normal((3, 4, 6, 2)),
         "/CTDSE/outputs/PopInt": rng.random(7).astype(np.float32),
         "/CTDSE/inputs/kz_step": np.int32(2),
         "/CTDSE/inputs/Nx_max": np.int64(8000),
         "/CTDSE/inputs/dx": 0.1,
         "/global_inputs/gas_preset": b"Ar",
         "/empty_1d": np.zeros(0),
         "/top": np.arange(12, dtype=np.int32).reshape(3, 4)}
    for i in range(21):  # three symbol-table nodes of eight entries
        d["/many/d%02d" % i] = np.full(i + 1, i, dtype=np.int64)
    attrs = {"/CTDSE/outputs/SourceTerm": {"units": b"[a.u.]"}, "/many": {"note": "grouped", "count": np.int32(21)},
             "/top": {"scale": np.array([1.5, 2.5])}}
    return d, attrs


def _same(got, want):
    if isinstance(want, (bytes, str)):
        return np.asarray(got).reshape(()).item() == (want.encode() if isinstance(want, str) else want)
    want = np.asarray(want)
    return got.shape == want.shape and got.dtype == want.dtype and np.array_equal(got, want)


def test_write_read_round_trip(tmp_path):
    d, attrs = _sample()
    p = str(tmp_path / "t.h5")
    h5lite.write(p, d, attrs)
    data, got_attrs = h5lite.read_all(p)
    assert set(data) == set(d)
    for k, v in d.items():
        assert _same(data[k], v), k
    assert got_attrs["/CTDSE/outputs/SourceTerm"]["units"] == b"[a.u.]"
    assert got_attrs["/many"]["note"] == b"grouped" and int(got_attrs["/many"]["count"]) == 21
    np.testing.assert_array_equal(got_attrs["/top"]["scale"], [1.5, 2.5])
    f = h5lite.H5File(p)
    assert f.keys("/many") == sorted("d%02d" % i for i in range(21))
    assert "/CTDSE/inputs/dx" in f and "/CTDSE/nothing" not in f
    with pytest.raises(ValueError):
        h5lite.write(p, {"/a": 1.0, "/a/b": 2.0})
    with pytest.raises(NotImplementedError):
        h5lite.write(p, {"/c": np.zeros(3, dtype=np.complex128)})


@pytest.mark.skipif(H5DUMP is None, reason="HDF5 command-line tools not installed")
def test_hdf5_library_reads_the_written_file(tmp_path):
    d, attrs = _sample()
    p = str(tmp_path / "t.h5")
    h5lite.write(p, d, attrs)
    listing = subprocess.run([H5DUMP, "-n", p], capture_output=True, text=True, check=True).stdout
    for k in d:
        assert "dataset    " + k in listing, k
    for k, v in d.items():
        if isinstance(v, bytes) or np.asarray(v).size == 0:
            continue
        raw = str(tmp_path / "raw.bin")
        subprocess.run([H5DUMP, "-d", k, "-b", "LE", "-o", raw, p], capture_output=True, check=True)
        a = np.asarray(v)
        np.testing.assert_array_equal(np.fromfile(raw, dtype=a.dtype.newbyteorder("<")), a.reshape(-1))
    text = subprocess.run([H5DUMP, "-d", "/global_inputs/gas_preset", p], capture_output=True, text=True,
                          check=True).stdout
    assert '"Ar"' in text and "DATASPACE  SCALAR" in text
    text = subprocess.run([H5DUMP, "-A", "-d", "/CTDSE/outputs/SourceTerm", p], capture_output=True, text=True,
                          check=True).stdout
    assert 'ATTRIBUTE "units"' in text and '"[a.u.]"' in text


@pytest.mark.skipif(H5REPACK is None, reason="HDF5 command-line tools not installed")
def test_reads_a_file_written_by_the_hdf5_library(tmp_path):
    """h5repack rewrites the file with the HDF5 library's own object placement (contiguous
    layout kept), and h5lite reads the library's file back."""
    d, attrs = _sample()
    p, q = str(tmp_path / "t.h5"), str(tmp_path / "repacked.h5")
    h5lite.write(p, d, attrs)
    subprocess.run([H5REPACK, "-l", "CONTI", p, q], capture_output=True, check=True)
    data, got_attrs = h5lite.read_all(q)
    for k, v in d.items():
        assert _same(data[k], v), k
    assert got_attrs["/CTDSE/outputs/SourceTerm"]["units"] == b"[a.u.]"


def _archive_inputs(preset=True, trg_a=None, eguess=None):
    d = {"/CTDSE/inputs/Nx_max": np.int32(800), "/CTDSE/inputs/dx": 0.1, "/CTDSE/inputs/InterpByDTorNT": np.int32(0),
         "/CTDSE/inputs/dt": 0.25, "/CTDSE/inputs/Ntinterp": np.int32(1), "/CTDSE/inputs/x_int": 2.0,
         "/CTDSE/inputs/CV_criterion_of_GS": 1e-25, "/CTDSE/inputs/kz_step": np.int32(2),
         "/CTDSE/inputs/kr_step": np.int32(1), "/CTDSE/inputs/Nz_max": np.int32(4), "/CTDSE/inputs/Nr_max": np.int32(3)}
    if preset:
        d["/global_inputs/gas_preset"] = b"Kr"
    if trg_a is not None:
        d["/CTDSE/inputs/trg_a"] = trg_a
    if eguess is not None:
        d["/CTDSE/inputs/Eguess"] = eguess
    return d


def test_archive_inputs_follow_readinputs(tmp_path):
    p = str(tmp_path / "a.h5")
    h5lite.write(p, _archive_inputs())
    inp = farm.read_inputs(h5lite.H5File(p))
    assert inp["num_r"] == 1600 and inp["dx"] == 0.1 and inp["dt"] == 0.25 and inp["x_int"] == 2.0
    assert (inp["a"], inp["e_guess"]) == farm.GAS_PRESETS["Kr"] == (1.3676, -0.5145)
    assert (inp["kz_step"], inp["kr_step"], inp["Nz_max"], inp["Nr_max"]) == (2, 1, 4, 3)
    # trg_a and Eguess in the CTDSE inputs take precedence over the preset (tools_hdf5.c:106-131)
    h5lite.write(p, _archive_inputs(trg_a=1.2, eguess=-0.6))
    inp = farm.read_inputs(h5lite.H5File(p))
    assert (inp["a"], inp["e_guess"]) == (1.2, -0.6)
    h5lite.write(p, _archive_inputs(preset=False, trg_a=1.2))
    with pytest.raises(ValueError):  # the reference exits: no ground-state energy guess
        farm.read_inputs(h5lite.H5File(p))
    d = _archive_inputs()
    d["/global_inputs/gas_preset"] = b"Og"
    h5lite.write(p, d)
    with pytest.raises(ValueError):
        farm.read_inputs(h5lite.H5File(p))


def test_inputs_save_load_round_trip(tmp_path, dll):
    """inputs_def.save_to_hdf5 writes the reference's names and types (float32 'f', int32 'i');
    load_from_hdf5 reads them back into a fresh structure, arrays allocated so the library's
    destructor frees them.  A second save into the same file keeps the first values (the block
    stops at the first existing name, as in the reference)."""
    from mma_amd.ctdse.structures import inputs_def
    p = str(tmp_path / "in.h5")
    a = inputs_def()
    a.init_default_inputs(Eguess=-0.5792, num_r=400, dx=0.2, trg_a=1.1893, CV=1e-20, x_int=3.0, writewft=1, tprint=5.0)
    t = np.arange(50) * 0.25
    a.init_time_and_field(dll, t=t, E=0.05 * np.sin(0.057 * t))
    dll.init_GS(a)
    a.save_to_hdf5(p)
    f = h5lite.H5File(p)
    assert f["/TDSE_inputs/dx"].dtype == np.float32 and f["/TDSE_inputs/N_r_grid"].dtype == np.int32
    assert f["/TDSE_inputs/psi0"].shape == (401, 2) and f["/TDSE_inputs/Efield"].shape == (50,)
    b = inputs_def()
    b.load_from_hdf5(p)
    assert b.num_r == 400 and b.analy.writewft == 1 and b.Efield.Nt == 50
    assert b.dx == float(np.float32(0.2)) and b.trg.a == float(np.float32(1.1893))
    np.testing.assert_array_equal(b.get_GS(), a.get_GS().astype(np.complex64).astype(np.complex128))
    np.testing.assert_array_equal(b.get_Efield(), a.get_Efield().astype(np.float32).astype(np.float64))
    np.testing.assert_array_equal(b.get_xgrid(), a.get_xgrid().astype(np.float32).astype(np.float64))
    assert b.Einit == float(np.float32(a.Einit))
    c = inputs_def()
    c.init_default_inputs(dx=0.5)
    c.save_to_hdf5(p)
    assert h5lite.H5File(p)["/TDSE_inputs/dx"] == np.float32(0.2)
    for s in (a, b):
        s.delete(dll)


def test_outputs_save_load_round_trip(tmp_path, dll):
    """outputs_def.save_to_hdf5 / load_from_hdf5 on outputs filled from arrays (no run needed):
    float32 series, (Nomega, 2) spectra, the wavefunction rows as psi_re / psi_im."""
    from mma_amd.ctdse.structures import inputs_def, outputs_def, _c_doubles, _c_rows
    rng = np.random.default_rng(5)
    nt, nom, nx = 40, 21, 11
    o = outputs_def()
    o.Nt, o.Nomega = nt, nom
    series = {k: rng.standard_normal(nt) for k in ("Efield", "sourceterm", "PopTot", "PopInt", "expval")}
    series["tgrid"] = np.arange(nt) * 0.25
    for k, v in series.items():
        setattr(o, k, _c_doubles(v))
    o.omegagrid = _c_doubles(np.arange(nom) * 0.1)
    spec = {k: rng.standard_normal(2 * nom) for k in ("FEfield", "Fsourceterm")}
    for k, v in spec.items():
        setattr(o, k, _c_doubles(v))
    rows = rng.standard_normal((4, 2 * nx))
    o.psi = _c_rows(list(rows))
    inp = inputs_def()
    inp.init_default_inputs(num_r=nx - 1, writewft=1, tprint=2.5)  # 10 steps per row
    p = str(tmp_path / "out.h5")
    o.save_to_hdf5(p, inputs=inp)
    f = h5lite.H5File(p)
    assert f["/TDSE_outputs/Fsourceterm"].shape == (nom, 2) and f["/TDSE_outputs/psi_re"].shape == (4, nx)
    q = outputs_def()
    q.load_from_hdf5(p)
    assert q.Nt == nt and q.Nomega == nom
    f32 = lambda a: np.asarray(a, dtype=np.float32).astype(np.float64)  # noqa: E731
    np.testing.assert_array_equal(q.get_sourceterm(), f32(series["sourceterm"]))
    np.testing.assert_array_equal(q.get_tgrid(), f32(series["tgrid"]))
    fs = f32(spec["Fsourceterm"])
    np.testing.assert_array_equal(q.get_Fsourceterm(), fs[0::2] + 1j * fs[1::2])
    np.testing.assert_array_equal(q.get_wavefunction(inp, grids=False), o.get_wavefunction(inp, grids=False)
                                  .astype(np.complex64).astype(np.complex128))
    dll.free_mtrx(q.psi, 4)
    dll.free_mtrx(o.psi, 4)
    for x in (o, q):
        x.delete(dll)
