"""Host side of the poststackmig driver on the CPU (libpsm_host.so):

* every function include/psm_host.h declares is exported and bound;
* the configuration parser follows config_parser.cpp (defaults, quotes,
  comments, booleans, validation messages);
* SEG-Y ingest equals an independent numpy reading of the reference CLI's
  golden input files (IEEE and IBM, shuffled and missing traces);
* the IBM codec equals the numpy restatement and round-trips;
* the SEG-Y writer reproduces the reference CLI's output files byte for byte
  when given the reference's output samples (headers copied or synthesised,
  IBM encoding included);
* the constant and SEG-Y velocity providers equal the oracle's restatement;
* the CLI reports configuration errors as the reference does, exit status 1.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import velocity_oracle as V
from _golden import GOLDEN, _read_bytes, inputs, load, output

from poststackmig3d_amd import host_api, segy

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "psm_host.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(psmh_[a-z_0-9]+)\s*\(", text)))


def test_host_library_exports_every_declared_symbol():
    lib = host_api.load()
    names = _declared()
    assert len(names) >= 12
    bound = {s[0] for s in host_api.SIGNATURES}
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) <= bound, set(names) - bound


def _cfg(tmp_path, text):
    p = tmp_path / "config.txt"
    p.write_text(text)
    return str(p)


def test_config_defaults_quotes_and_comments(tmp_path):
    c = host_api.parse_config(_cfg(tmp_path, (
        "# comment\n\n  input_data = 'in file.sgy'  \noutput_data=\"out.sgy\"\n"
        "inline_step=25\ncrossline_step = 12.5\nvelocity=2000.0\n")))
    assert c["input_data"] == "in file.sgy" and c["output_data"] == "out.sgy"
    assert c["inline_step"] == 25.0 and c["crossline_step"] == 12.5
    assert c["velocity"] == "2000.0"
    assert (c["inline_padding"], c["crossline_padding"]) == (0, 0)
    assert c["angle_aperture"] == 30.0 and c["amp_correction"] == 1 and c["n_threads"] == 0


@pytest.mark.parametrize("val,expect", [("true", 1), ("ON", 1), ("yes", 1), ("1", 1),
                                        ("false", 0), ("Off", 0), ("no", 0), ("0", 0)])
def test_config_booleans(tmp_path, val, expect):
    c = host_api.parse_config(_cfg(tmp_path, (
        "input_data=a\noutput_data=b\ninline_step=1\ncrossline_step=1\nvelocity=v.txt\n"
        f"amp_correction={val}\ninline_padding=3\ncrossline_padding=4\nangle_aperture=45\n"
        "n_threads=2\n")))
    assert c["amp_correction"] == expect
    assert (c["inline_padding"], c["crossline_padding"], c["angle_aperture"], c["n_threads"]) == (3, 4, 45.0, 2)


@pytest.mark.parametrize("text,msg", [
    ("input_data=a\noutput_data=b\nvelocity=1\n", "Missing required parameters: inline_step, crossline_step"),
    ("input_data=a\noutput_data=b\ninline_step=1\ncrossline_step=1\nvelocity=1\nbad line\n",
     "Invalid format at line 6: expected key=value, got: bad line"),
    ("input_data=a\noutput_data=b\ninline_step=x\ncrossline_step=1\nvelocity=1\n",
     "Error parsing config values: Cannot parse double value: x"),
    ("input_data=a\noutput_data=b\ninline_step=1\ncrossline_step=1\nvelocity=1\namp_correction=maybe\n",
     "Error parsing config values: Cannot parse boolean value: maybe"),
    ("input_data=a\noutput_data=b\ninline_step=-1\ncrossline_step=1\nvelocity=1\n",
     "inline_step must be positive"),
    ("input_data=a\noutput_data=b\ninline_step=1\ncrossline_step=1\nvelocity=1\nangle_aperture=90\n",
     "angle_aperture must be in range (0, 90)"),
    ("input_data=a\noutput_data=b\ninline_step=1\ncrossline_step=1\nvelocity=1\ninline_padding=-2\n",
     "inline_padding must be non-negative"),
])
def test_config_errors(tmp_path, text, msg):
    with pytest.raises(host_api.PsmError, match=re.escape(msg)):
        host_api.parse_config(_cfg(tmp_path, text))


def test_config_missing_file(tmp_path):
    with pytest.raises(host_api.PsmError, match="Config file not found"):
        host_api.parse_config(str(tmp_path / "nope.txt"))


def _write_bytes(tmp_path, arr, name):
    p = tmp_path / name
    p.write_bytes(arr.tobytes())
    return str(p)


@pytest.mark.parametrize("path", GOLDEN, ids=lambda p: p.split("/")[-1][:-4])
def test_ingest_matches_numpy_reader(tmp_path, path):
    fx = load(path)
    inp = _write_bytes(tmp_path, fx["input_file"], "in.sgy")
    seg = segy.read(inp)
    cube_np, ill, xll = segy.to_cube(seg)
    with host_api.SegyFile(inp) as f:
        assert f.n_traces == seg["traces"].shape[0]
        assert (f.n_t, f.dt_us, f.fmt) == (seg["n_t"], seg["dt_us"], seg["fmt"])
        np.testing.assert_array_equal(f.inline_labels, ill)
        np.testing.assert_array_equal(f.crossline_labels, xll)
        cube = f.read_cube()
    np.testing.assert_array_equal(cube, cube_np)


def test_ingest_duplicates_zero_labels_and_errors(tmp_path):
    n_t = 7
    tr = np.arange(5 * n_t, dtype=np.float32).reshape(5, n_t)
    il = np.array([3, 3, 0, 5, 3])
    xl = np.array([1, 2, 9, 1, 1])      # (3,1) appears twice: the later trace wins
    p = str(tmp_path / "d.sgy")
    segy.write(p, tr, il, xl)
    with host_api.SegyFile(p) as f:
        np.testing.assert_array_equal(f.inline_labels, [3, 5])
        np.testing.assert_array_equal(f.crossline_labels, [1, 2])
        c = f.read_cube()
    np.testing.assert_array_equal(c[0, 0], tr[4])
    np.testing.assert_array_equal(c[0, 1], tr[1])
    np.testing.assert_array_equal(c[1, 0], tr[3])
    np.testing.assert_array_equal(c[1, 1], np.zeros(n_t, np.float32))
    bad = str(tmp_path / "bad.sgy")
    open(bad, "wb").write(b"\0" * 3600)
    with pytest.raises(host_api.PsmError, match="Sample interval"):
        host_api.SegyFile(bad)
    with pytest.raises(host_api.PsmError, match="Cannot open SEGY file"):
        host_api.SegyFile(str(tmp_path / "missing.sgy"))


def test_ibm_codec():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.normal(0, 1, 4000), rng.normal(0, 1e-20, 500), rng.normal(0, 1e20, 500),
                        [0.0, -0.0, 1.0, -1.0, 1 / 16, 15.999999, 3.4e38, 1e-38]]).astype(np.float32)
    ibm = host_api.ieee_to_ibm(x)
    np.testing.assert_array_equal(ibm, segy.ieee_to_ibm(x))
    np.testing.assert_array_equal(host_api.ibm_to_ieee(ibm), segy.ibm_to_ieee(ibm))
    # a value decoded from IBM encodes back to the same word
    u = rng.integers(0, 2 ** 32, 4000, dtype=np.uint64).astype(np.uint32)
    u = u[((u >> 24) & 0x7F) > 40]           # stay inside the float range
    u = u[((u >> 24) & 0x7F) < 90]
    u = u[(u & 0x00F00000) != 0]             # normalised fractions
    back = host_api.ieee_to_ibm(host_api.ibm_to_ieee(u))
    np.testing.assert_array_equal(back, u)


@pytest.mark.parametrize("path", GOLDEN, ids=lambda p: p.split("/")[-1][:-4])
def test_writer_reproduces_reference_output_bytes(tmp_path, path):
    """Given the reference's own output samples, the writer must produce the
    reference CLI's output file byte for byte."""
    fx = load(path)
    inp = _write_bytes(tmp_path, fx["input_file"], "in.sgy")
    img, oil, oxl = output(fx)
    out = str(tmp_path / "out.sgy")
    with host_api.SegyFile(inp) as f:
        f.write(out, oil, oxl, img)
    got = np.fromfile(out, np.uint8)
    ref = fx["output_file"]
    assert got.size == ref.size
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("path", [p for p in GOLDEN if "table" not in p],
                         ids=lambda p: p.split("/")[-1][:-4])
def test_velocity_providers_match_oracle(tmp_path, path):
    fx = load(path)
    cube, vel_ref, ill, xll, dt, num, vmax, seg = inputs(fx)
    mode = str(fx["vel_mode"])
    if mode == "segy":
        param = _write_bytes(tmp_path, fx["vel_file"], "vel.sgy")
    else:
        param = str(fx["velocity_param"])
    got = host_api.velocity_cube(param, ill, xll, cube.shape[2], dt)
    np.testing.assert_allclose(got, vel_ref, rtol=2e-7, atol=0)


def test_velocity_segy_interpolation_paths(tmp_path):
    """Velocity SEG-Y on another grid: missing traces, labels between and
    outside the velocity grid, a coarser time axis."""
    rng = np.random.default_rng(7)
    vc = (1500 + 1000 * rng.random((4, 5, 30))).astype(np.float32)
    vil = np.array([10, 12, 14, 20])
    vxl = np.array([1, 3, 5, 7, 9])
    p = str(tmp_path / "v.segy")
    segy.write_cube(p, vc, vil, vxl, dt_us=4000)
    seg = segy.read(p)
    cube, a, b = segy.to_cube(seg)
    prov = V.SegyVelocity(cube, a, b, 0.004)
    ill = np.array([9, 10, 11, 12, 16, 21, 30])
    xll = np.array([0, 1, 2, 5, 8, 9, 12])
    ref = V.velocity_cube(prov, ill, xll, 55, 0.002)
    got = host_api.velocity_cube(p, ill, xll, 55, 0.002)
    np.testing.assert_allclose(got, ref, rtol=2e-7, atol=0)


def test_constant_velocity_errors():
    with pytest.raises(host_api.PsmError, match="Velocity must be positive, got: -5.000000"):
        host_api.velocity_cube("-5", [1], [1], 3, 0.002)
    with pytest.raises(host_api.PsmError, match="Cannot open velocity table file"):
        host_api.velocity_cube("2e3", [1], [1], 3, 0.002)


def test_cli_usage_and_config_errors(tmp_path):
    cli = host_api.CLI_PATH
    r = subprocess.run([cli], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage:" in r.stderr
    r = subprocess.run([cli, str(tmp_path / "none.txt")], capture_output=True, text=True)
    assert r.returncode == 1 and "Error: Config file not found" in r.stderr
    cfg = _cfg(tmp_path, "input_data=a\noutput_data=b\n")
    r = subprocess.run([cli, cfg], capture_output=True, text=True)
    assert r.returncode == 1
    assert "Error: Missing required parameters: inline_step, crossline_step, velocity" in r.stderr
