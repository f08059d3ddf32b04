"""The TransducerModel driver on the DeviceSolver, from a job .vtu to solution and table
files (serial names, and the partitioned .pvtu/.pvtt layout): every velocity table and the
last solution equal the oracle run of the same job, bit for bit, with the step counts of the
reference's write-time schedule."""
import numpy as np
import pytest

from oracle import oracle_rsg as orc
from pannde_amd import cases, rsg, transducer_model, vtkio, water_column_model

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("partitioned", [False, True])
def test_job_file_to_outputs_matches_oracle(tmp_path, monkeypatch, partitioned):
    src = cases.al_block(seed=41, nx=11, ny=9, nz=6, nsteps=40, n_sensors=1)
    wt = np.array([0.0, 7.3, 15.5, 30.2, 38.0]) * src.dt  # uneven intervals, one at t = 0
    job = tmp_path / "job.vtu"
    vtkio.case_to_job(src, str(job), wt, fmt="appended", compress=True)
    run_dir = tmp_path / "run"
    run_dir.mkdir()
    monkeypatch.chdir(run_dir)  # the tables go to the working directory, as the reference writes them
    out = str(tmp_path / "solution")
    r = transducer_model.run(str(job), out, rsg.RSG_FP64, partitioned=partitioned)
    case, _ = vtkio.job_case(str(job))
    steps = vtkio.schedule_steps(case.dt, wt)
    assert r["steps"] == steps[-1] and steps[0] == 0
    V, S, t, tr = orc.run_oracle(case, nsteps=int(steps[-1]))
    if partitioned:
        last = vtkio.read_pvtu("%s_%d.pvtu" % (out, len(wt) - 1))
        assert (tmp_path / ("solution_%d_0.vtu" % (len(wt) - 1))).exists()
        table, ext = vtkio.read_pvtt, ".pvtt"
    else:
        last = vtkio.read_vtu("%s_%d.vtu" % (out, len(wt) - 1))
        table, ext = vtkio.read_vtt, ".vtt"
    for k, name in enumerate(rsg.STATE_V):
        assert np.array_equal(last.point_data[name], V[k]), name
    for k, name in enumerate(rsg.STATE_S):
        assert np.array_equal(last.cell_data[name], S[k]), name
    assert last.field_data["time"][0] == t
    coords = table(str(run_dir / ("node_coords" + ext)))
    assert np.array_equal(coords["node_x"], case.x[case.sensors]) and case.sensors.shape[0] > 0
    for k, n in enumerate(steps):
        vel = table(str(run_dir / ("velocities_%d%s" % (k, ext))))
        ref = tr[n - 1] if n else np.zeros((case.sensors.shape[0], 3))
        for c, name in enumerate(("Vx", "Vy", "Vz")):
            assert np.array_equal(vel[name], ref[:, c]), (k, name)


def test_water_column_job_matches_oracle(tmp_path):
    """WaterColumnModel counterpart: two water columns (XX, YY, ZZ moments of the HostUTModel
    form); every written solution equals the oracle's at its step count."""
    from test_vtkio import two_column_case

    src = two_column_case(nsteps=30)
    wt = np.array([0.0, 9.5, 21.2, 29.0]) * src.dt
    job = tmp_path / "wc.vtu"
    vtkio.case_to_water_column_job(src, str(job), wt, fmt="binary", compress=True)
    out = str(tmp_path / "wc_out")
    r = water_column_model.run(str(job), out, rsg.RSG_FP64)
    case, _ = vtkio.water_column_case(str(job))
    steps = vtkio.schedule_steps(case.dt, wt)
    assert r["steps"] == steps[-1]
    for k, n in enumerate(steps):
        sol = vtkio.read_vtu("%s_%d.vtu" % (out, k))
        V, S, t, _ = orc.run_oracle(case, nsteps=int(n))
        for j, name in enumerate(rsg.STATE_V):
            assert np.array_equal(sol.point_data[name], V[j]), (k, name)
        for j, name in enumerate(rsg.STATE_S):
            assert np.array_equal(sol.cell_data[name], S[j]), (k, name)
        assert sol.field_data["time"][0] == t
    assert np.abs(sol.cell_data["Sxx"]).max() > 0
