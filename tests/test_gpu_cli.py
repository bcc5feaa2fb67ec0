"""The simulate CLI: plain (per-function) and fused (engine) modes agree with the oracle."""
import numpy as np
import pytest

from helpers import oracle_sim, rel_inf

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

ARGS = ["--elements", "Fe,Mn,Ni,Co,Cu", "--temperature", "873K", "--resolution", "0.025", "--c11", "240",
        "--c12", "150", "--c44", "120", "--molar_volume", "7.1e-6", "--system_dim", "2", "--mw", "3", "--ppw", "8",
        "--direction", "0,0,1", "--initial_composition", "0.1,0.2,0.2,0.3", "--kappa_value", "5e-16",
        "--synthetic_table", "--seed", "3", "--total_timesteps", "30", "--no_save_outputs"]


def test_cli_plain_and_fused_match_oracle():
    from phasefield5d_amd.problem import build_problem
    from phasefield5d_amd.simulate import run
    from phasefield5d_amd.solver.config import parse_args_to_config
    plain = run(ARGS + ["--mode", "plain"])
    fused = run(ARGS + ["--mode", "fused", "--precision", "fp64"])
    f32 = run(ARGS + ["--mode", "fused", "--precision", "fp32"])
    pb = build_problem(parse_args_to_config(ARGS))
    ora = oracle_sim(pb)
    X, dt, t = pb.X_init.copy(), pb.dt0, 0.0
    for _ in range(30):
        X, _, dt, t, _ = ora.step(X, dt, t)
    assert plain["steps"] == fused["steps"] == 30
    assert rel_inf(plain["X"] - pb.X_init, X - pb.X_init) < 1e-10
    assert rel_inf(fused["X"] - pb.X_init, X - pb.X_init) < 1e-10
    assert abs(plain["time"] - t) <= 1e-12 * t and abs(fused["time"] - t) <= 1e-12 * t
    assert np.abs(f32["X"] - X).max() < 1e-6
    # the same save steps in both modes, with the same diagnostics
    tp, tf = plain["traces"], fused["traces"]
    assert tp["timestep"] and tp["timestep"] == tf["timestep"] == f32["traces"]["timestep"]
    for k in ("time", "dt", "composition_change_max", "comp_min_s0", "comp_max_s3", "comp_var_s1", "comp_fe_min",
              "f_grad"):
        assert np.allclose(tp[k], tf[k], rtol=1e-9, atol=0), k
    assert np.allclose(np.array(tp["mass"]), np.array(tf["mass"]), rtol=1e-12)
    assert np.allclose(np.array(tp["net"]), np.array(tf["net"]), rtol=1e-6,
                       atol=1e-9 * np.abs(np.array(tp["net"])).max())


def test_cli_run_directory_layout(tmp_path):
    """The reference's run layout (solver/io.py): metadata, data/step_*.npz with the
    reference keys, the trace table with the reference columns."""
    import json
    import os

    import pandas as pd

    from phasefield5d_amd.simulate import run
    from phasefield5d_amd.solver.io import init_traces
    args = [a for a in ARGS if a != "--no_save_outputs"] + ["--total_timesteps", "12", "--results_root",
                                                              str(tmp_path), "--mode", "fused"]
    out = run(args)
    d = out["run_dir"]
    assert d.startswith(str(tmp_path)) and "Fe20Mn10Ni20Co20Cu30_at_873K_2dim" in d
    assert "elastic_cubic_normal001" in d and os.path.basename(d).startswith("cells24x24_dx")
    meta = json.load(open(os.path.join(d, "metadata.json")))
    assert meta["number_of_cells"] == [24, 24] and meta["total_timesteps"] == 12
    files = sorted(os.listdir(os.path.join(d, "data")))
    assert files[0] == "step_000000000.npz" and "step_000000000_initial.npz" in files
    assert "step_000000011.npz" in files
    z = np.load(os.path.join(d, "data", "step_000000011.npz"))
    assert sorted(z.files) == ["current_composition", "dt", "time", "timestep"]
    assert int(z["timestep"]) == 11 and np.allclose(z["current_composition"], out["X"], rtol=0, atol=0)
    df = pd.read_csv(os.path.join(d, "timeseries_info.csv"))
    assert list(df.columns) == list(init_traces(4).keys())
    assert df["timestep"].iloc[0] == 0 and df["timestep"].iloc[-1] == 11
    js = json.load(open(os.path.join(d, "timeseries_info.json")))
    assert js["rows"] == len(df)
