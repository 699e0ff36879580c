"""Host-side failures of the drop-in executables, forced on the CPU (no GPU is
touched: every case fails before a device context is created).  The reference
checks none of these (SOAP_integration.cu:508-518 reads without checking, a
missing file crashes or yields garbage); here each ends with a message on stderr
naming the problem and a non-zero exit status: wrong argument counts, missing and
short .bin files, a wavelength grid that is neither linear nor log-uniform, an
output length beyond the vector, missing CCF text files, and a configuration file
that is missing, lacks a key or holds a non-number."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from soap_gpu_amd.build import BIN


def tool(name):
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        pytest.fail(f"{path} missing: run __graft_entry__.build() first")
    return path


def run(name, args, cwd):
    r = subprocess.run([tool(name), *[str(a) for a in args]], cwd=str(cwd), capture_output=True, text=True,
                       timeout=60)
    return r.returncode, r.stderr


def _init_args(d, n=4096, n_mu=3, wave="wave.bin", cube="cube.bin", mub="mub.bin", no_block=8, no_thread=512):
    # SOAP_initialize.cu:513-537 order
    return [n, 25.0, 0.0, 90.0, 696000.0, 0.29, 0.34, 32, 1107.0, 0.005, 0.005, 50.0, no_block, no_thread,
            d / cube, str(d) + "/out_", n_mu, d / mub, d / wave, 0.2, -0.1, 0.0, 0.0]


def _write(path, a):
    np.asarray(a, dtype=np.float64).tofile(path)


def test_initialize_failures(tmp_path):
    n, n_mu = 4096, 3
    rc, err = run("SOAP_initialize", [], tmp_path)
    assert rc != 0 and "usage" in err
    rc, err = run("SOAP_initialize", _init_args(tmp_path)[:-1], tmp_path)
    assert rc != 0 and "usage" in err
    rc, err = run("SOAP_initialize", _init_args(tmp_path), tmp_path)  # no files at all
    assert rc != 0 and "cannot open" in err and "wave.bin" in err
    wave = 5000.0 + 0.005 * np.arange(n)
    _write(tmp_path / "wave.bin", wave[: n // 2])  # short
    rc, err = run("SOAP_initialize", _init_args(tmp_path), tmp_path)
    assert rc != 0 and "expected 4096 values" in err
    _write(tmp_path / "wave.bin", wave)
    _write(tmp_path / "cube.bin", np.ones(n * n_mu - 1))  # one value short
    rc, err = run("SOAP_initialize", _init_args(tmp_path), tmp_path)
    assert rc != 0 and "cube.bin" in err
    _write(tmp_path / "cube.bin", np.ones(n * n_mu))
    _write(tmp_path / "mub.bin", [0.9, 0.5])
    bad = wave.copy()
    bad[100:] += 0.001 * np.arange(n - 100) ** 1.5  # neither linear nor log-uniform
    _write(tmp_path / "wave.bin", bad)
    rc, err = run("SOAP_initialize", _init_args(tmp_path), tmp_path)
    assert rc != 0 and "neither linear nor log-uniform" in err
    _write(tmp_path / "wave.bin", wave)
    rc, err = run("SOAP_initialize", _init_args(tmp_path, no_block=9), tmp_path)  # 9 x 512 > 4096
    assert rc != 0 and "exceeds the vector size" in err


def test_integration_and_resolution_failures(tmp_path):
    for name in ("SOAP_integration", "SOAP_integration_solar"):
        rc, err = run(name, ["1", "2"], tmp_path)
        assert rc != 0 and "usage" in err
    args = [32, 20, 2, 5778, 663, 90.0, 250, tmp_path / "phases.txt", tmp_path / "fac.bin", 4096,
            str(tmp_path) + "/", str(tmp_path) + "/", 8, 512, tmp_path / "q.bin", tmp_path / "sp.bin", "x", 0.005,
            3, tmp_path / "wave.bin", tmp_path / "qs.bin", tmp_path / "sps.bin", tmp_path / "facs.bin",
            str(tmp_path) + "/"]
    rc, err = run("SOAP_integration", args, tmp_path)
    assert rc != 0 and "cannot open" in err and "wave.bin" in err
    rc, err = run("SOAP_resolution", [], tmp_path)
    assert rc != 0 and "usage" in err
    res = [2, 4096, 1107.0, str(tmp_path) + "/", str(tmp_path) + "/", 1, 4, "x", str(tmp_path) + "/q"]
    rc, err = run("SOAP_resolution", res, tmp_path)  # 1 x 4 pixels
    assert rc != 0 and "[8, N]" in err
    res[5], res[6] = 8, 512
    rc, err = run("SOAP_resolution", res, tmp_path)
    assert rc != 0 and "cannot open" in err and "_quiet_sun_wave.bin" in err


def test_bis_fit2_failures(tmp_path):
    rc, err = run("BIS_FIT2", [], tmp_path)
    assert rc != 0 and "usage" in err
    rc, err = run("BIS_FIT2", [5], tmp_path)
    assert rc != 0 and "CCF_data.txt" in err
    np.savetxt(tmp_path / "CCF_data.txt", np.array([np.arange(5.0), np.ones(5)]).T)
    rc, err = run("BIS_FIT2", [5], tmp_path)
    assert rc != 0 and "CCF_error.txt" in err
    np.savetxt(tmp_path / "CCF_error.txt", np.ones(3))
    rc, err = run("BIS_FIT2", [5], tmp_path)
    assert rc != 0 and "shorter" in err


def test_config_driver_failures(tmp_path):
    rc, err = run("soap_run", [tmp_path / "missing.cfg"], tmp_path)
    assert rc != 0 and "cannot open config" in err
    (tmp_path / "a.cfg").write_text("[main]\ngrid = 32\n")
    rc, err = run("soap_run", [tmp_path / "a.cfg"], tmp_path)
    assert rc != 0 and "config: missing [" in err
    (tmp_path / "b.cfg").write_text("[main]\ngrid = thirty\n")
    rc, err = run("soap_run", [tmp_path / "b.cfg"], tmp_path)
    assert rc != 0 and ("not an integer" in err or "missing" in err)
