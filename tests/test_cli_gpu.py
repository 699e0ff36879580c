"""Drop-in host side on the GPU: the config-driven driver (soap_run, in place of
run_all_SOAP.py) and the stage executables with the reference's argv
(SOAP_initialize, SOAP_integration, SOAP_resolution, BIS_FIT2), run on a
temporary directory laid out like the reference's, checked against the oracle
pipeline and against each other.

Reference behaviour followed: run_all_SOAP.py:56-145 (configuration and derived
parameters), :148-220 (phase locator and region files, 'real' keyword),
:224-250 (stage argv), :259-412 (CCF, fit, RV referencing, .npz keys),
:443-449 (RV summary)."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

import oracle as O
import pipeline_ref as PR
from soap_gpu_amd import synth
from soap_gpu_amd.build import BIN

pytestmark = pytest.mark.gpu

T_STAR = 5778
CUBE_FILES = {  # run_all_SOAP.py:94-100
    "q": f"new_noconti_convec_T_eff_{T_STAR}.bin",
    "qsed": f"new_conti_convec_T_eff_{T_STAR}.bin",
    "sp": f"new_noconti_convec_T_eff_{T_STAR - 663}.bin",
    "spsed": f"new_conti_T_eff_{T_STAR - 663}.bin",
    "fac": f"new_noconti_convec_T_eff_{T_STAR + 250}.bin",
    "facsed": f"new_conti_T_eff_{T_STAR + 250}.bin",
}
REGIONS = [(30.0, 10.0, 0.15, 1.0), (32.0, 12.0, 0.08, 0.0)]  # a facula, then a spot inside it
NO_THREAD = 500


def tool(name):
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        pytest.fail(f"{path} missing: run __graft_entry__.build() first")
    return path


def run(cmd, cwd):
    print("run:", " ".join(str(c) for c in cmd), flush=True)
    r = subprocess.run([str(c) for c in cmd], cwd=cwd, capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        raise AssertionError(f"{cmd[0]} exited {r.returncode}\nstdout:\n{r.stdout[-2000:]}\nstderr:\n{r.stderr[-2000:]}")
    return r.stdout


def make_workload(tables, grid=150):
    w = synth.config0(grid=grid)
    w.phases = np.arange(0.0, 0.2, 0.04)  # rot_num 0.2, rot_step 0.04
    w.regions = [list(REGIONS) for _ in w.phases]
    w.star["lb"] = O.lb_params(T_STAR, tables["LB_G2"], tables["LB_K0"], tables["LB_M0"], tables["Temp_array"])
    return w


def write_inputs(root, w, tables):
    inp = os.path.join(root, "input_spec")
    os.makedirs(inp, exist_ok=True)
    os.makedirs(os.path.join(root, "input_mask"), exist_ok=True)
    w.wave.tofile(os.path.join(inp, "wave.bin"))
    np.asarray(w.mu_bounds, dtype=np.float64).tofile(os.path.join(inp, "mu_boundary_array.bin"))
    for k, name in CUBE_FILES.items():
        np.ascontiguousarray(w.cubes[k], dtype=np.float64).tofile(os.path.join(inp, name))
    np.savez(os.path.join(inp, "Limb_brightening_info.npz"),
             **{k: np.asarray(tables[k], dtype=np.float64) for k in ("LB_G2", "LB_K0", "LB_M0", "Temp_array")})
    np.savetxt(os.path.join(root, "input_mask", "G2_mask.txt"), np.stack([w.mask_l, w.mask_w], axis=1))


def _f(x):
    return repr(float(x))


def write_config(root, w, keyword="nreal", regions=REGIONS):
    s = w.star
    lines = ["[main]", f"grid : {s['grid']}", "nrho : 20", "instrument_reso : 115000", "rot_num: 0.2",
             "rot_step: 0.04", "", "[star]", "radius_sun : 696000", "radius : 1.0", f"prot : {_f(s['prot'])}",
             f"W_ratio : {_f(s['w_ratio'])}", f"I : {_f(s['incl'])}", f"Tstar : {T_STAR}", "Tdiff_spot : 663",
             "Tdiff_faculae : 250", f"limb1 : {_f(s['u1'])}", f"limb2 : {_f(s['u2'])}", f"N_act: {len(regions)}", ""]
    for i, (lat, lon, size, ty) in enumerate(regions):
        lines += [f"[active_regions_{i + 1}]", f"longitude: {_f(lon)}", f"latitude: {_f(lat)}", f"size: {_f(size)}",
                  f"type: {int(ty)}", ""]
    lines += ["[data_io]", f"no_thread: {NO_THREAD}", f"input_prefix: {root}", f"output_prefix: {root}/out",
              "file_prefix: test", "input_dir: /input_spec/", "maps_dir: /maps/", "raw_dir: /raw/",
              "spec_dir: /spec/", "mu_array_name: mu_boundary_array.bin", "wave_array_name: wave.bin",
              "CCF_window: 20000", "CCF_size: 500.", "mask_prefix: /input_mask/", f"keyword: {keyword}"]
    path = os.path.join(root, "config_hpc.cfg")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


def out_prefix(root):
    return os.path.join(root, f"out_phoenix{T_STAR}")


@pytest.fixture(scope="module")
def driver_run(tmp_path_factory, tables):
    root = str(tmp_path_factory.mktemp("soaprun"))
    w = make_workload(tables)
    write_inputs(root, w, tables)
    cfg = write_config(root, w)
    run([tool("soap_run"), cfg], cwd=root)
    ref = PR.run_oracle(w)
    return root, w, ref


def load_bin(path, dtype=np.float64):
    return np.fromfile(path, dtype=dtype)


def test_driver_npz_products(driver_run):
    root, w, ref = driver_run
    op = out_prefix(root)
    P = len(w.phases)
    names = {"tot": 2, "flux": 0, "bconv": 1}
    summary = np.load(os.path.join(op, "test_GPU_SOAP_RV_data.npz"))
    assert sorted(summary.files) == ["bconv", "flux", "time", "tot"]
    np.testing.assert_array_equal(summary["time"], w.phases)
    for name, t in names.items():
        d = np.load(os.path.join(op, f"test_RV_{name}_GPU.npz"))
        assert sorted(d.files) == sorted(["time", "RV", "ccf", "model", "continuum", "contrast", "span", "fwhm",
                                          "depth", "bis"])
        assert d["ccf"].shape == (P + 1, len(w.vgrid)) and d["depth"].shape == (P + 1, 91)
        np.testing.assert_array_equal(d["RV"], summary[name])
        np.testing.assert_allclose(d["RV"], ref["rv"][:, t], rtol=0, atol=1e-2)  # 1 cm/s
        np.testing.assert_allclose(d["fwhm"], ref["par"][:, t, 4], rtol=0, atol=1e-2)
        np.testing.assert_allclose(d["span"], ref["par"][:, t, 2], rtol=0, atol=1e-2)
        np.testing.assert_allclose(d["ccf"], ref["ccf"][:, t], rtol=0, atol=1e-10)
        np.testing.assert_allclose(d["bis"], ref["bis"][:, t], rtol=0, atol=1e-2)
    # the spot and facula produce a measurable RV signal
    assert np.ptp(summary["tot"]) > 0.1


def test_driver_stage_files(driver_run):
    root, w, ref = driver_run
    op = out_prefix(root)
    G = w.star["grid"]
    n = w.n
    q = load_bin(os.path.join(op, "test_quiet_sun_spec.bin"))
    qs = load_bin(os.path.join(op, "test_SED_quiet_sun_spec.bin"))
    assert q.size == n and qs.size == n
    assert PR.normwise(q, ref["quiet"][0]) <= 1e-12 and PR.normwise(qs, ref["quiet"][1]) <= 1e-12
    np.testing.assert_array_equal(load_bin(os.path.join(op, "test_quiet_sun_wave.bin")), w.wave)
    rm = load_bin(os.path.join(op, "testregion_map.bin"), np.int32)
    np.testing.assert_array_equal(rm, ref["maps"]["region"].ravel())
    lower_q = load_bin(os.path.join(op, "test_quiet_sun_lower_spec.bin"))
    assert PR.normwise(lower_q, ref["lower"][-1, 1]) <= 1e-12
    tn = ("flux", "bconv", "tot")
    for i in range(len(w.phases)):
        for t in range(3):
            raw = load_bin(os.path.join(op, "raw", f"test_Spec_{tn[t]}_{i}_.bin"))
            assert PR.normwise(raw, ref["raw"][i, t]) <= 1e-12, (i, t)
            low = load_bin(os.path.join(op, "spec", f"test_lower_spec_{tn[t]}_{i}_.bin"))
            assert PR.normwise(low, ref["lower"][i, t]) <= 1e-12, (i, t)
        tm = load_bin(os.path.join(op, "raw", f"test_temp_map_{i}_.bin"))
        bm = load_bin(os.path.join(op, "raw", f"test_spot_binary_map_{i}_.bin"))
        np.testing.assert_array_equal(tm, ref["phase_maps"][i][0].ravel())
        np.testing.assert_array_equal(bm, ref["phase_maps"][i][1].ravel())
        assert tm.size == G * G
    # region files of the generated ('nreal') run
    info = np.loadtxt(os.path.join(op, "maps", "final_info.txt"))
    np.testing.assert_array_equal(info, [[i, len(REGIONS)] for i in range(len(w.phases))])
    m0 = np.loadtxt(os.path.join(op, "maps", "maps0_for_all.txt"))
    np.testing.assert_array_equal(m0, np.array(REGIONS))
    np.testing.assert_array_equal(np.loadtxt(os.path.join(op, "test_phase_locator.txt")), w.phases)
    assert os.path.exists(os.path.join(op, "test_config_file.cfg"))


def test_stage_executables_match_driver(driver_run, tmp_path):
    """The four stage executables, with the argv of run_all_SOAP.py:224-250, reproduce
    the driver's files."""
    root, w, ref = driver_run
    s = w.star
    op = out_prefix(root)
    inp = os.path.join(root, "input_spec") + "/"
    o2 = str(tmp_path) + "/"
    for d in ("raw", "spec"):
        os.makedirs(o2 + d, exist_ok=True)
    n = w.n
    no_block = n // NO_THREAD
    sig = synth.sigma_inst_ms(115000)
    step = round(w.wave[3] - w.wave[2], 5)
    lb = w.star["lb"]
    n_mu = w.n_mu
    common = [n, s["prot"], s["w_ratio"], s["incl"], 696000.0, s["u1"], s["u2"], s["grid"], sig, step, step, 50.0,
              no_block, NO_THREAD]
    tail = [n_mu, inp + "mu_boundary_array.bin", inp + "wave.bin", lb[0], lb[1], lb[2], lb[3]]
    run([tool("SOAP_initialize"), *common, inp + CUBE_FILES["q"], o2 + "test", *tail], cwd=o2)
    run([tool("SOAP_initialize"), *common, inp + CUBE_FILES["qsed"], o2 + "test_SED", *tail], cwd=o2)
    maps_dir = os.path.join(op, "maps") + "/"
    run([tool("SOAP_integration"), s["grid"], 20, len(w.phases), T_STAR, 663, s["incl"], 250,
         os.path.join(op, "test_phase_locator.txt"), inp + CUBE_FILES["fac"], n, o2 + "raw/", maps_dir, no_block,
         NO_THREAD, inp + CUBE_FILES["q"], inp + CUBE_FILES["sp"], "test", step, n_mu, inp + "wave.bin",
         inp + CUBE_FILES["qsed"], inp + CUBE_FILES["spsed"], inp + CUBE_FILES["facsed"], o2], cwd=o2)
    run([tool("SOAP_resolution"), len(w.phases), n, sig, o2 + "raw/", o2 + "spec/", no_block, NO_THREAD, "test",
         o2 + "test"], cwd=o2)
    for name in ("mu_map", "delta_map", "limb_map", "LB_map"):
        np.testing.assert_array_equal(load_bin(o2 + f"test{name}.bin"), load_bin(os.path.join(op, f"test{name}.bin")))
    np.testing.assert_array_equal(load_bin(o2 + "testregion_map.bin", np.int32),
                                  load_bin(os.path.join(op, "testregion_map.bin"), np.int32))
    for name in ("_quiet_sun_spec", "_SED_quiet_sun_spec", "_quiet_sun_lower_spec"):
        a = load_bin(o2 + "test" + name + ".bin")
        b = load_bin(os.path.join(op, "test" + name + ".bin"))
        assert PR.normwise(a, b) <= 1e-13, name
    tn = ("flux", "bconv", "tot")
    for i in range(len(w.phases)):
        for t in range(3):
            a = load_bin(o2 + f"raw/test_Spec_{tn[t]}_{i}_.bin")
            np.testing.assert_array_equal(a, load_bin(os.path.join(op, "raw", f"test_Spec_{tn[t]}_{i}_.bin")))
            a = load_bin(o2 + f"spec/test_lower_spec_{tn[t]}_{i}_.bin")
            b = load_bin(os.path.join(op, "spec", f"test_lower_spec_{tn[t]}_{i}_.bin"))
            assert PR.normwise(a, b) <= 1e-12, (i, t)
            assert PR.normwise(a, ref["lower"][i, t]) <= 1e-12, (i, t)
        np.testing.assert_array_equal(load_bin(o2 + f"raw/test_temp_map_{i}_.bin"),
                                      load_bin(os.path.join(op, "raw", f"test_temp_map_{i}_.bin")))


def test_integration_solar_executable(driver_run, tmp_path):
    """SOAP_integration_solar (same argv) gives the oracle's solar-mode deltas."""
    root, w, ref = driver_run
    s = w.star
    op = out_prefix(root)
    inp = os.path.join(root, "input_spec") + "/"
    o2 = str(tmp_path) + "/"
    os.makedirs(o2 + "raw", exist_ok=True)
    n = w.n
    step = round(w.wave[3] - w.wave[2], 5)
    run([tool("SOAP_integration_solar"), s["grid"], 20, len(w.phases), T_STAR, 663, s["incl"], 250,
         os.path.join(op, "test_phase_locator.txt"), inp + CUBE_FILES["fac"], n, o2 + "raw/",
         os.path.join(op, "maps") + "/", n // NO_THREAD, NO_THREAD, inp + CUBE_FILES["q"], inp + CUBE_FILES["sp"],
         "test", step, w.n_mu, inp + "wave.bin", inp + CUBE_FILES["qsed"], inp + CUBE_FILES["spsed"],
         inp + CUBE_FILES["facsed"], op + "/"], cwd=o2)
    a = w.act
    tn = ("flux", "bconv", "tot")
    for i in range(len(w.phases)):
        temp, bi = ref["phase_maps"][i]
        exp = O.integrate_phase(O.GRID_LOG, w.wave, ref["maps"], temp, bi, w.cubes, s["grid"], a["t_star"],
                                a["t_diff_fac"], solar_mode=1)
        for t in range(3):
            got = load_bin(o2 + f"raw/test_Spec_{tn[t]}_{i}_.bin")
            assert PR.normwise(got, exp[t]) <= 1e-12, (i, t)


def test_bis_fit2_executable(driver_run, tmp_path):
    """BIS_FIT2 on CCF_data.txt / CCF_error.txt (as run_all_SOAP.py:372-381 writes
    them) returns the driver's fit of the same CCF."""
    root, w, ref = driver_run
    d = np.load(os.path.join(out_prefix(root), "test_RV_tot_GPU.npz"))
    i = 2
    np.savetxt(tmp_path / "CCF_data.txt", np.array([w.vgrid, d["ccf"][i]]).T)
    np.savetxt(tmp_path / "CCF_error.txt", np.ones(w.vgrid.size).T)
    run([tool("BIS_FIT2"), w.vgrid.size], cwd=str(tmp_path))
    par = load_bin(str(tmp_path / "ccf_parameter.bin"))
    rv_abs = d["RV"][i] + ref["par"][-1, 2, 3]  # undo the referencing to the quiet frame
    np.testing.assert_allclose(par[[0, 1]], [d["continuum"][i], d["contrast"][i]], rtol=1e-12)
    assert abs(par[3] - rv_abs) <= 1e-2
    assert abs(par[4] - d["fwhm"][i]) <= 1e-6 and abs(par[2] - d["span"][i]) <= 1e-6
    np.testing.assert_allclose(load_bin(str(tmp_path / "ccf_model.bin")), d["model"][i], rtol=0, atol=1e-12)
    np.testing.assert_allclose(load_bin(str(tmp_path / "ccf_bis.bin")), d["bis"][i], rtol=0, atol=1e-6)


def test_bis_fit2_executable_with_errors(driver_run, tmp_path):
    """BIS_FIT2 with a non-unit CCF_error.txt (BIS_FIT2.cpp:114-119): the fit
    weights residuals by 1/err, as the oracle's fit with the same errors."""
    root, w, ref = driver_run
    d = np.load(os.path.join(out_prefix(root), "test_RV_tot_GPU.npz"))
    y = d["ccf"][1]
    err = 1e-3 * (1 + (w.vgrid / 8000.0) ** 2) * np.random.default_rng(7).uniform(0.5, 2.0, w.vgrid.size)
    np.savetxt(tmp_path / "CCF_data.txt", np.array([w.vgrid, y]).T)
    np.savetxt(tmp_path / "CCF_error.txt", err.T)
    run([tool("BIS_FIT2"), w.vgrid.size], cwd=str(tmp_path))
    par = load_bin(str(tmp_path / "ccf_parameter.bin"))
    v_txt = np.loadtxt(tmp_path / "CCF_data.txt")
    ro = O.fit_ccf(v_txt[:, 0], v_txt[:, 1], np.loadtxt(tmp_path / "CCF_error.txt"))
    assert abs(par[3] - ro["par"][3]) <= 1e-4 and abs(par[4] - ro["par"][4]) <= 1e-4
    assert abs(par[2] - ro["par"][2]) <= 1e-4
    np.testing.assert_allclose(par[:2], ro["par"][:2], rtol=1e-7)  # both fits stop at xtol 1e-8


def test_bis_fit2_executable_bad_inputs(tmp_path):
    """Host-side failures of the BIS_FIT2 shim end with a message and a non-zero
    status before any GPU work: wrong argc, missing files, short files, and a
    zero error (the reference would divide by it)."""
    exe = tool("BIS_FIT2")

    def status(args):
        r = subprocess.run([exe, *[str(a) for a in args]], cwd=str(tmp_path), capture_output=True, text=True,
                           timeout=60)
        return r.returncode, r.stderr

    assert status([])[0] != 0
    assert status([5, 6])[0] != 0
    rc, err = status([5])
    assert rc != 0 and "CCF_data.txt" in err
    np.savetxt(tmp_path / "CCF_data.txt", np.array([np.arange(5.0), np.ones(5)]).T)
    np.savetxt(tmp_path / "CCF_error.txt", np.ones(3))
    rc, err = status([5])
    assert rc != 0 and "shorter" in err
    np.savetxt(tmp_path / "CCF_error.txt", np.array([1, 1, 0.0, 1, 1]))
    rc, err = status([5])
    assert rc != 0 and "non-zero" in err


def test_driver_real_keyword_ragged_regions(tmp_path, tables):
    """keyword 'real': phases and per-phase region lists of different lengths (one
    phase with none) are read from the files instead of the configuration."""
    root = str(tmp_path)
    w = make_workload(tables, grid=96)
    w.phases = np.array([0.0, 0.1, 0.25, 0.3])
    w.regions = [[REGIONS[0]], [], list(REGIONS), [(-20.0, 90.0, 0.12, 0.0), REGIONS[0], (5.0, 40.0, 0.1, 0.0)]]
    write_inputs(root, w, tables)
    cfg = write_config(root, w, keyword="real")
    op = out_prefix(root)
    os.makedirs(os.path.join(op, "maps"))
    np.savetxt(os.path.join(op, "test_phase_locator.txt"), w.phases)
    ids = [3, 1, 0, 2]  # map files are found through final_info.txt
    np.savetxt(os.path.join(op, "maps", "final_info.txt"), [[ids[i], len(r)] for i, r in enumerate(w.regions)])
    for i, r in enumerate(w.regions):
        np.savetxt(os.path.join(op, "maps", f"maps{ids[i]}_for_all.txt"), np.array(r).reshape(-1, 4))
    run([tool("soap_run"), cfg, "--no-maps"], cwd=root)
    ref = PR.run_oracle(w)
    tn = ("flux", "bconv", "tot")
    for i in range(len(w.phases)):
        for t in range(3):
            raw = load_bin(os.path.join(op, "raw", f"test_Spec_{tn[t]}_{ids[i]}_.bin"))
            if not w.regions[i]:
                assert np.all(raw == 0.0)
            else:
                assert PR.normwise(raw, ref["raw"][i, t]) <= 1e-12, (i, t)
    for name, t in (("tot", 2), ("flux", 0), ("bconv", 1)):
        d = np.load(os.path.join(op, f"test_RV_{name}_GPU.npz"))
        np.testing.assert_allclose(d["RV"], ref["rv"][:, t], rtol=0, atol=1e-2)
        assert d["RV"][1] == pytest.approx(0.0, abs=1e-6)  # no active region: the quiet frame


def test_driver_size_evolution(tmp_path, tables):
    """--size-evolution: the spot-size evolution of run_all_SOAP_evo.py:214-220, size
    at phase p = sqrt(size^2 + 0.002 (p - p0) prot) for the default rate, written to
    the per-phase map files and used for the spectra (checked against the oracle
    run with the evolved regions)."""
    root = str(tmp_path)
    w = make_workload(tables, grid=96)
    write_inputs(root, w, tables)
    cfg = write_config(root, w)
    run([tool("soap_run"), cfg, "--no-maps", "--size-evolution"], cwd=root)
    op = out_prefix(root)
    prot = float(w.star["prot"])
    evolved = []
    for i, ph in enumerate(w.phases):
        t = (ph - w.phases[0]) * prot
        regs = [(lat, lon, float(np.sqrt(size * size + 0.002 * t)), ty) for lat, lon, size, ty in REGIONS]
        evolved.append(regs)
        got = np.loadtxt(os.path.join(op, "maps", f"maps{i}_for_all.txt")).reshape(-1, 4)
        np.testing.assert_allclose(got, np.array(regs), rtol=1e-15, atol=0)
    assert evolved[-1][0][2] > REGIONS[0][2] + 1e-3  # the sizes do change over the run
    w.regions = evolved
    ref = PR.run_oracle(w)
    tn = ("flux", "bconv", "tot")
    for i in range(len(w.phases)):
        for t in range(3):
            raw = load_bin(os.path.join(op, "raw", f"test_Spec_{tn[t]}_{i}_.bin"))
            assert PR.normwise(raw, ref["raw"][i, t]) <= 1e-12, (i, t)
