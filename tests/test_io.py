"""OVF 2.0 I/O (host code, runs without a GPU): core/test/test_io.cpp's round trips for every
format, chains, appends, and -- where the reference's own OVF library has been compiled into
oracle/_ref -- files byte-identical to the reference's writer and readable by its parser
(and the reference's files readable by the engine)."""
import ctypes
import json
import os

import numpy as np
import pytest

from spirit_amd import configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_goldens.json")))
REF_OVF = os.path.join(ROOT, "oracle", "_ref", "libovf_ref.so")

# (format, max abs error of a unit-vector round trip): binary exact (the reference's
# double-precision writer turns bin4 into Binary 8 too), text and CSV 12 decimals
FORMATS = {"bin": (0, 0.0), "bin4": (1, 0.0), "bin8": (2, 0.0), "text": (3, 1e-12), "csv": (4, 1e-12)}


@pytest.fixture
def make_state(tmp_path):
    from spirit_amd import state
    made = []

    def make(cfg):
        path = str(tmp_path / ("input%d.cfg" % len(made)))
        configs.write_config(cfg, path)
        p = state.setup(path, True)
        assert p
        made.append(p)
        return p

    yield make
    for p in made:
        state.delete(p)


def spins(p, idx_image=-1):
    from spirit_amd import system
    return np.array(system.get_spin_directions(p, idx_image=idx_image))


def set_spins(p, s, idx_image=-1):
    from spirit_amd import system
    system.get_spin_directions(p, idx_image=idx_image)[:] = s


def unit(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1)[:, None]


@pytest.mark.parametrize("fmt", list(FORMATS))
def test_image_roundtrip(make_state, tmp_path, fmt):
    """test_io.cpp:22-70: minus z written, plus z set, the file read back gives minus z; and a
    random configuration to the format's precision."""
    from spirit_amd import configuration, io
    code, tol = FORMATS[fmt]
    p = make_state(GOLD["fd_pairs"]["cfg"])
    fn = str(tmp_path / ("image_%s.ovf" % fmt))
    configuration.minus_z(p)
    io.image_write(p, fn, code, "io test")
    configuration.plus_z(p)
    io.image_read(p, fn)
    s = spins(p)
    assert s.shape == (4, 3)
    np.testing.assert_allclose(s, np.tile([0.0, 0.0, -1.0], (4, 1)), rtol=0, atol=1e-12)
    p2 = make_state(configs.scaled(configs.SQUARE_64, [5, 4, 3]))
    s0 = unit(np.random.default_rng(7), 60)
    set_spins(p2, s0)
    io.image_write(p2, fn, code)
    configuration.plus_z(p2)
    io.image_read(p2, fn)
    # the reader renormalises each spin
    ref = s0 / np.linalg.norm(s0, axis=1)[:, None]
    assert np.abs(spins(p2) - ref).max() <= max(tol, 1e-15)
    assert io.n_images_in_file(p2, fn) == 1


def test_append_and_chain(make_state, tmp_path):
    """Appended segments are counted in the header; a chain written to a file is read back into
    a one-image chain, which grows to the file's images (test_io.cpp:104-170)."""
    from spirit_amd import chain, configuration, io
    cfg = configs.scaled(configs.SQUARE_64, [4, 4, 1])
    p = make_state(cfg)
    rng = np.random.default_rng(3)
    fn = str(tmp_path / "appended.ovf")
    written = []
    for k in range(3):
        s = unit(rng, 16)
        set_spins(p, s)
        written.append(s)
        io.image_append(p, fn, 2, "image %d" % k)
    assert io.n_images_in_file(p, fn) == 3
    for k in range(3):
        io.image_read(p, fn, idx_image_infile=k)
        np.testing.assert_allclose(spins(p), written[k], rtol=0, atol=1e-15)
    # a chain of three images: minus z, random, plus z
    chain.image_to_clipboard(p)
    chain.insert_image_before(p)
    chain.insert_image_before(p)
    chain.jump_to_image(p, 0)
    configuration.minus_z(p)
    chain.jump_to_image(p, 1)
    s_rand = unit(rng, 16)
    set_spins(p, s_rand)
    chain.jump_to_image(p, 2)
    configuration.plus_z(p)
    for fmt, (code, tol) in FORMATS.items():
        fc = str(tmp_path / ("chain_%s.ovf" % fmt))
        io.chain_write(p, fc, code, "chain")
        assert io.n_images_in_file(p, fc) == 3
        q = make_state(cfg)
        io.chain_read(q, fc)
        assert chain.get_noi(q) == 3
        assert np.abs(spins(q, 0) - [0.0, 0.0, -1.0]).max() <= 1e-12
        ref = s_rand / np.linalg.norm(s_rand, axis=1)[:, None]
        assert np.abs(spins(q, 1) - ref).max() <= max(tol, 1e-15)
        assert np.abs(spins(q, 2) - [0.0, 0.0, 1.0]).max() <= 1e-12
    # appending the chain to an existing file adds its images
    io.chain_append(p, fn, 2, "chain")
    assert io.n_images_in_file(p, fn) == 6


def test_not_ovf(make_state, tmp_path):
    from spirit_amd import io
    p = make_state(configs.scaled(configs.SQUARE_64, [2, 2, 1]))
    fn = str(tmp_path / "plain.txt")
    open(fn, "w").write("1 0 0\n0 1 0\n")
    assert io.n_images_in_file(p, fn) == -1
    before = spins(p).copy()
    io.image_read(p, fn)  # logged error, spins unchanged
    np.testing.assert_array_equal(spins(p), before)
    io.image_append(p, fn, 2)  # refuses to append to a non-OVF file
    assert open(fn).read() == "1 0 0\n0 1 0\n"


# ---------------------------------------------------------------------------
# the reference's OVF library (core/thirdparty/ovf), compiled from its sources
# ---------------------------------------------------------------------------
class OvfSegment(ctypes.Structure):
    _fields_ = [("title", ctypes.c_char_p), ("comment", ctypes.c_char_p), ("valuedim", ctypes.c_int),
                ("valueunits", ctypes.c_char_p), ("valuelabels", ctypes.c_char_p), ("meshtype", ctypes.c_char_p),
                ("meshunit", ctypes.c_char_p), ("pointcount", ctypes.c_int), ("n_cells", ctypes.c_int * 3),
                ("N", ctypes.c_int), ("step_size", ctypes.c_float * 3), ("bounds_min", ctypes.c_float * 3),
                ("bounds_max", ctypes.c_float * 3), ("lattice_constant", ctypes.c_float),
                ("origin", ctypes.c_float * 3)]


class OvfFile(ctypes.Structure):
    _fields_ = [("file_name", ctypes.c_char_p), ("version", ctypes.c_int), ("found", ctypes.c_bool),
                ("is_ovf", ctypes.c_bool), ("n_segments", ctypes.c_int), ("_state", ctypes.c_void_p)]


def ref_ovf():
    if not os.path.exists(REF_OVF):
        pytest.skip("oracle/_ref/libovf_ref.so not built (the reference checkout is absent)")
    L = ctypes.CDLL(REF_OVF)
    P = ctypes.POINTER
    L.ovf_open.argtypes = [ctypes.c_char_p]
    L.ovf_open.restype = P(OvfFile)
    L.ovf_segment_create.restype = P(OvfSegment)
    for name in ("ovf_write_segment_8", "ovf_append_segment_8"):
        getattr(L, name).argtypes = [P(OvfFile), P(OvfSegment), P(ctypes.c_double), ctypes.c_int]
        getattr(L, name).restype = ctypes.c_int
    L.ovf_read_segment_header.argtypes = [P(OvfFile), ctypes.c_int, P(OvfSegment)]
    L.ovf_read_segment_header.restype = ctypes.c_int
    L.ovf_read_segment_data_8.argtypes = [P(OvfFile), ctypes.c_int, P(OvfSegment), P(ctypes.c_double)]
    L.ovf_read_segment_data_8.restype = ctypes.c_int
    L.ovf_close.argtypes = [P(OvfFile)]
    return L


def ref_segment(L, p, comment):
    """The segment IO_Image_Write fills (IO.cpp:293-360, OVF_File.cpp:12-39)."""
    from spirit_amd import geometry, system
    from spirit_amd import spiritlib
    seg = L.ovf_segment_create()
    s = seg.contents
    title = b"SPIRIT Version " + spiritlib.load_spirit_library().Spirit_Version_Full()
    s.title, s.comment, s.valuedim = title, comment, 3
    s.valuelabels, s.valueunits = b"spin_x spin_y spin_z", b"none none none"
    s.meshtype, s.meshunit = b"rectangular", b"nm"
    nc = geometry.get_n_cells(p)
    s.n_cells[0], s.n_cells[1], s.n_cells[2] = nc[0] * geometry.get_n_cell_atoms(p), nc[1], nc[2]
    s.N = system.get_nos(p)
    lo, hi = geometry.get_bounds(p)
    for k in range(3):
        s.bounds_min[k], s.bounds_max[k] = lo[k] * 0.1, hi[k] * 0.1
        s.origin[k] = 0.0
        s.step_size[k] = 0.1  # lattice constant 1 Angstrom, simple cubic
    return seg


def test_reference_writer_bytes(make_state, tmp_path):
    """Every format: the engine's file equals, byte for byte, the file the reference's OVF
    library writes for the same segment and data; appended files too."""
    from spirit_amd import io, spiritlib
    spiritlib.load_spirit_library().Spirit_Version_Full.restype = ctypes.c_char_p
    L = ref_ovf()
    p = make_state(configs.scaled(configs.SQUARE_64, [5, 4, 3]))
    s0 = unit(np.random.default_rng(11), 60)
    set_spins(p, s0)
    data = np.ascontiguousarray(s0, dtype=np.float64)
    dp = data.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    for fmt, (code, _) in FORMATS.items():
        ours, theirs = str(tmp_path / ("ours_%s.ovf" % fmt)), str(tmp_path / ("ref_%s.ovf" % fmt))
        io.image_write(p, ours, code, "bytes test")
        io.image_append(p, ours, code, "second")
        f = L.ovf_open(theirs.encode())
        assert L.ovf_write_segment_8(f, ref_segment(L, p, b"bytes test"), dp, code) == -1
        L.ovf_close(f)
        f = L.ovf_open(theirs.encode())
        assert L.ovf_append_segment_8(f, ref_segment(L, p, b"second"), dp, code) == -1
        L.ovf_close(f)
        assert open(ours, "rb").read() == open(theirs, "rb").read(), fmt


def test_reference_reader_and_writer_interoperate(make_state, tmp_path):
    """The reference's parser reads the engine's chain file (segment count, header, data), and
    the engine reads a chain file the reference's library wrote."""
    from spirit_amd import chain, io, spiritlib
    spiritlib.load_spirit_library().Spirit_Version_Full.restype = ctypes.c_char_p
    L = ref_ovf()
    cfg = configs.scaled(configs.SQUARE_64, [4, 3, 2])
    p = make_state(cfg)
    rng = np.random.default_rng(5)
    chain.image_to_clipboard(p)
    chain.insert_image_after(p)
    chain.insert_image_after(p)
    imgs = [unit(rng, 24) for _ in range(3)]
    for k in range(3):
        set_spins(p, imgs[k], idx_image=k)
    fn = str(tmp_path / "engine_chain.ovf")
    io.chain_write(p, fn, 2, "interop")
    f = L.ovf_open(fn.encode())
    assert f.contents.is_ovf and f.contents.n_segments == 3
    for k in range(3):
        seg = L.ovf_segment_create()
        assert L.ovf_read_segment_header(f, k, seg) == -1
        assert seg.contents.N == 24 and seg.contents.valuedim == 3
        assert seg.contents.comment.decode() == "Image %d of 3. interop" % (k + 1)
        out = np.zeros((24, 3))
        assert L.ovf_read_segment_data_8(f, k, seg, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == -1
        np.testing.assert_array_equal(out, imgs[k])
    L.ovf_close(f)
    # the reference writes, the engine reads
    fr = str(tmp_path / "reference_chain.ovf")
    for k in range(3):
        f = L.ovf_open(fr.encode())
        d = np.ascontiguousarray(imgs[2 - k])
        fn_w = L.ovf_write_segment_8 if k == 0 else L.ovf_append_segment_8
        assert fn_w(f, ref_segment(L, p, b"ref"), d.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 3) == -1
        L.ovf_close(f)
    q = make_state(cfg)
    io.chain_read(q, fr)
    assert chain.get_noi(q) == 3
    for k in range(3):
        ref = imgs[2 - k] / np.linalg.norm(imgs[2 - k], axis=1)[:, None]
        assert np.abs(spins(q, k) - ref).max() <= 1e-12


@pytest.mark.gpu
def test_gpu_run_files_match_reference_writer(make_state, tmp_path):
    """After an LLG run on the GPU (Depondt, 50 steps, spins back from the device), every
    format the engine writes equals, byte for byte, what the reference's OVF library writes
    for the same segment and spins, and the reference's parser reads the engine's file back
    to the same values."""
    from spirit_amd import io, simulation, spiritlib
    print("CASE ovf bytes after a gpu run", flush=True)
    spiritlib.load_spirit_library().Spirit_Version_Full.restype = ctypes.c_char_p
    L = ref_ovf()
    p = make_state(configs.scaled(configs.SQUARE_64, [9, 7, 2]))
    set_spins(p, unit(np.random.default_rng(12), 126))
    simulation.start(p, simulation.METHOD_LLG, 2, n_iterations=50)
    s1 = spins(p)
    data = np.ascontiguousarray(s1, dtype=np.float64)
    dp = data.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    for fmt, (code, _) in FORMATS.items():
        ours, theirs = str(tmp_path / ("gpu_%s.ovf" % fmt)), str(tmp_path / ("gpu_ref_%s.ovf" % fmt))
        io.image_write(p, ours, code, "after a gpu run")
        f = L.ovf_open(theirs.encode())
        assert L.ovf_write_segment_8(f, ref_segment(L, p, b"after a gpu run"), dp, code) == -1
        L.ovf_close(f)
        assert open(ours, "rb").read() == open(theirs, "rb").read(), fmt
        f = L.ovf_open(ours.encode())
        seg = L.ovf_segment_create()
        assert L.ovf_read_segment_header(f, 0, seg) == -1
        out = np.zeros((126, 3))
        assert L.ovf_read_segment_data_8(f, 0, seg, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == -1
        L.ovf_close(f)
        assert np.abs(out - s1).max() <= FORMATS[fmt][1], fmt


@pytest.mark.gpu
def test_llg_configuration_output_files(tmp_path):
    """Method_LLG::Save_Current (Method_LLG.cpp:310-495) as Method::Iterate calls it
    (Method.cpp:57-113): <tag>_Image-00_Spins-initial.ovf, -final.ovf, _<iteration>.ovf per
    save and -archive.ovf with one segment per save.  The loop saves after the body that
    starts at iteration 10 and 20 (so those files hold the spins after 11 and 21 steps), the
    initial and final saves also write the step and archive outputs: archive segments 0, 10,
    20, 30.  The energy files (_Energy-initial/-final/_<iteration>.txt, -archive.txt) carry
    the header and one line per save, energies per spin to 10 decimals.  Spins are checked
    against the oracle's run (binary files; reading renormalises, so 1e-15 against the run's
    own final spins)."""
    import oracle as O
    from spirit_amd import io, simulation, state, system
    from spirit_amd.parameters import llg
    print("CASE llg output files", flush=True)
    out = tmp_path / "out"
    out.mkdir()
    cfg = dict(bravais_lattice="sc", n_basis_cells=[8, 6, 1], boundary_conditions=[1, 1, 0], mu_s=[2.0],
               external_field_magnitude=1.0, external_field_normal=[0, 0, 1], anisotropy_magnitude=0.1,
               anisotropy_normal=[0, 0, 1], n_shells_exchange=1, jij=[5.0], ddi_method="none", llg_dt=1e-3,
               llg_damping=0.3, llg_temperature=0.0, llg_force_convergence=1e-14)
    text = configs.config_text(cfg) + "\n".join([
        "output_file_tag run", "llg_output_folder %s" % out, "llg_output_any 1", "llg_output_initial 1",
        "llg_output_final 1", "llg_output_configuration_step 1", "llg_output_configuration_archive 0",
        "llg_output_configuration_filetype 0", "llg_output_energy_step 1", "llg_output_energy_archive 1"]) + "\n"
    path = tmp_path / "input.cfg"
    path.write_text(text)
    p = state.setup(str(path), True)
    assert p
    try:
        assert llg.get_output_configuration(p) == (True, False, 0)
        llg.set_output_configuration(p, True, True, 0)
        rng = np.random.default_rng(3)
        s0 = unit(rng, 48)
        set_spins(p, s0)
        simulation.start(p, simulation.METHOD_LLG, simulation.SOLVER_DEPONDT, n_iterations=30, n_iterations_log=10)
        s_end = spins(p).copy()
        E_end = system.get_energy(p)
        names = sorted(f.name for f in out.iterdir())
        want = ["run_Image-00_Energy-archive.txt", "run_Image-00_Energy-final.txt", "run_Image-00_Energy-initial.txt",
                "run_Image-00_Energy_0.txt", "run_Image-00_Energy_10.txt", "run_Image-00_Energy_20.txt",
                "run_Image-00_Energy_30.txt",
                "run_Image-00_Spins-archive.ovf", "run_Image-00_Spins-final.ovf", "run_Image-00_Spins-initial.ovf",
                "run_Image-00_Spins_0.ovf", "run_Image-00_Spins_10.ovf", "run_Image-00_Spins_20.ovf",
                "run_Image-00_Spins_30.ovf"]
        assert names == want
        assert io.n_images_in_file(p, str(out / "run_Image-00_Spins-archive.ovf")) == 4
        io.image_read(p, str(out / "run_Image-00_Spins-archive.ovf"), 0)
        np.testing.assert_allclose(spins(p), s0, rtol=0, atol=1e-15)
        io.image_read(p, str(out / "run_Image-00_Spins-initial.ovf"))
        np.testing.assert_allclose(spins(p), s0, rtol=0, atol=1e-15)
        for f in ("run_Image-00_Spins-final.ovf", "run_Image-00_Spins_30.ovf"):
            io.image_read(p, str(out / f))
            np.testing.assert_allclose(spins(p), s_end, rtol=0, atol=1e-15)
        io.image_read(p, str(out / "run_Image-00_Spins-archive.ovf"), 3)
        np.testing.assert_allclose(spins(p), s_end, rtol=0, atol=1e-15)
        sysO = O.System(cfg)
        for label, k in (("10", 11), ("20", 21)):
            ref, _, _, _ = sysO.llg(s0, O.SOLVER_DEPONDT, k, dt=1e-3, damping=0.3)
            io.image_read(p, str(out / ("run_Image-00_Spins_%s.ovf" % label)))
            assert np.abs(spins(p) - ref).max() < 1e-9, label
            io.image_read(p, str(out / "run_Image-00_Spins-archive.ovf"), int(label) // 10)
            assert np.abs(spins(p) - ref).max() < 1e-9, label
        head = open(out / "run_Image-00_Spins_20.ovf", "rb").read(2000).decode("latin-1")
        assert "# Desc: LLG simulation (Depondt solver)" in head and "# Desc:      Iteration: 20" in head
        # energy files: Write_Energy_Header / Append_Image_Energy without readability lines
        arch = (out / "run_Image-00_Energy-archive.txt").read_text().splitlines()
        assert len(arch) == 5
        assert arch[0].split() == ["iteration", "E_tot"]
        assert [int(l.split()[0]) for l in arch[1:]] == [0, 10, 20, 30]
        last = arch[-1].split()
        assert len(last[1].split(".")[1]) == 10
        assert abs(float(last[1]) - E_end / 48) < 1e-6 * max(1.0, abs(E_end / 48))
        assert (out / "run_Image-00_Energy_20.txt").read_text().splitlines()[1].split()[0] == "20"
    finally:
        state.delete(p)
