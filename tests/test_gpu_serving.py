"""Serving loop on the GPU: jobs that arrive while others run, admission
against an HBM budget, per-job failure isolation; every finished job is
bit-identical to the same scene stepped alone and to the oracle."""
import numpy as np
import pytest

from fluidsolver_amd import abi, scenes
from fluidsolver_amd.collection import ParticleCollection
from fluidsolver_amd.serving import Server
from fluidsolver_amd.solver import Scene, memory_estimate

pytestmark = pytest.mark.gpu


def _jobs():
    """Uneven dam scenes (both solvers, different viscosities and step counts)."""
    out = []
    for q, (n, steps, solver, nu) in enumerate([(5, 4, 0, 5.0), (8, 6, 0, 0.05), (6, 3, 1, 5.0), (4, 7, 0, 1.0),
                                                (7, 5, 1, 0.5), (9, 2, 0, 5.0)]):
        c = scenes.dam_scene((n, n + 1, n), (3 * n, 2 * n, 2 * n), 3, seed=70 + q)
        d = abi.default_desc(3, solver)
        if solver == abi.SOLVER_IISPH:
            d.iisph_viscosity = nu
        else:
            d.sesph_viscosity = nu
            d.sesph_stiffness_k = 1000.0
        out.append((d, c, steps))
    return out


def _alone(ctx, d, c, steps):
    sc = Scene(ctx, d, c.copy())
    sc.step(steps)
    res = sc.download(c.copy())
    sc.close()
    return res


@pytest.mark.parametrize("serial,graph", [(True, False), (False, False), (False, True)])
def test_server_mix_equals_alone(gpu_context, oracle_lib, serial, graph):
    """Jobs arriving while others run.  Serial ticks admit, step and retire one
    after another.  Overlapped ticks (the default) step the running jobs on a
    worker thread while arrivals are uploaded and finished jobs downloaded on a
    second stream, so arrivals take their first step in the tick after their
    admission.  With graphs, they are switched on as the arrivals come, so the
    step's launches are captured (new graph keys) while the arrivals are
    allocated and uploaded, and again while the first finished job is
    downloaded.  Every job equals running it alone."""
    jobs = _jobs()
    print("case serving mix", len(jobs), "jobs", "serial" if serial else "overlapped", "graph" if graph else "",
          flush=True)
    try:
        srv = Server(gpu_context, serial_ticks=serial)
        ids = [srv.submit(d, c.copy(), s) for d, c, s in jobs[:3]]
        assert srv.tick() == 3
        assert srv.tick() == 3
        ids += [srv.submit(d, c.copy(), s) for d, c, s in jobs[3:]]  # arrivals while others run
        gpu_context.set_graph(graph)
        running = []
        while True:
            st = srv.stats()
            if st["jobs_queued"] == 0 and st["jobs_running"] == 0:
                break
            running.append(srv.tick())
        # the first jobs have 4, 6 and 3 steps: serial ticks step the arrivals
        # beside all three in tick 3; overlapped ticks from tick 4, beside the two
        # with steps left
        assert max(running) == (6 if serial else 5)
        assert srv.stats()["jobs_failed"] == 0
        for jid, (d, c, steps) in zip(ids, jobs):
            info = srv.job(jid)
            assert info.state == abi.JOB_DONE and info.steps_done == steps
            got = srv.collect(jid)
            ref = _alone(gpu_context, d, c, steps)
            np.testing.assert_array_equal(got.movement, ref.movement)
            np.testing.assert_array_equal(got.particle_data, ref.particle_data)
        srv.close()
    finally:
        gpu_context.set_graph(False)


def test_cancel_between_last_step_and_retirement(gpu_context):
    """With overlapped ticks a job is retired in the tick after its last step.
    Cancelling it in between retires it DONE, with the result of running it
    alone; the job beside it is unaffected."""
    jobs = _jobs()
    d, c, steps = jobs[5]  # 2 steps
    d1, c1, s1 = jobs[1]  # 6 steps: the second tick steps a running job, so it overlaps
    srv = Server(gpu_context)
    a = srv.submit(d1, c1.copy(), s1)
    b = srv.submit(d, c.copy(), steps)
    print("case cancel after the last step", flush=True)
    for _ in range(steps):
        assert srv.tick() == 2
    info = srv.job(b)
    assert info.state == abi.JOB_RUNNING and info.steps_done == steps
    srv.cancel(b)
    assert srv.job(b).state == abi.JOB_DONE
    np.testing.assert_array_equal(srv.collect(b).movement, _alone(gpu_context, d, c, steps).movement)
    while srv.stats()["jobs_running"] or srv.stats()["jobs_queued"]:
        srv.tick()
    np.testing.assert_array_equal(srv.collect(a).movement, _alone(gpu_context, d1, c1, s1).movement)
    assert srv.stats()["jobs_done"] == 2
    srv.close()


def test_server_slot_limit_and_cancel(gpu_context):
    jobs = _jobs()
    srv = Server(gpu_context, max_running=2)
    ids = [srv.submit(d, c.copy(), s) for d, c, s in jobs]
    assert srv.tick() == 2
    srv.cancel(ids[5])  # still queued
    srv.cancel(ids[0])  # running
    assert srv.job(ids[5]).state == abi.JOB_CANCELLED
    assert srv.job(ids[0]).state == abi.JOB_CANCELLED
    while srv.stats()["jobs_running"] or srv.stats()["jobs_queued"]:
        assert srv.tick() <= 2
    for q in (1, 2, 3, 4):
        d, c, steps = jobs[q]
        got = srv.collect(ids[q])
        np.testing.assert_array_equal(got.movement, _alone(gpu_context, d, c, steps).movement)
    srv.close()


def test_server_isolates_failed_job(gpu_context):
    """A job whose lattice overflows the 97-neighbor capacity fails alone; the
    jobs sharing its launches finish and match running alone."""
    jobs = _jobs()[:3]
    pos = (scenes.lattice((10, 10, 10)) * np.float32(0.45)).astype(np.float32)
    bad = ParticleCollection.from_arrays(3, pos, np.zeros(len(pos), np.uint8))
    srv = Server(gpu_context)
    ids = [srv.submit(d, c.copy(), s) for d, c, s in jobs]
    bad_id = srv.submit(abi.default_desc(3, abi.SOLVER_IISPH), bad, 3)
    while srv.stats()["jobs_running"] or srv.stats()["jobs_queued"]:
        srv.tick()
    info = srv.job(bad_id)
    assert info.state == abi.JOB_FAILED and info.status == abi.FSA_E_CAPACITY
    for jid, (d, c, steps) in zip(ids, jobs):
        np.testing.assert_array_equal(srv.collect(jid).movement, _alone(gpu_context, d, c, steps).movement)
    assert srv.stats()["jobs_failed"] == 1
    srv.close()


def test_server_isolates_job_leaving_the_key_range(gpu_context):
    """A job whose cells leave the 32-bit key range fails alone at the host-side set-up of
    its tick (before any launch); the jobs of the same batch take that step and finish
    equal to running alone.  The failed job is forgotten by collect without a copy."""
    jobs = _jobs()[:3]
    d, c, _ = jobs[0]
    far = ParticleCollection.from_arrays(3, np.array([[1500.0, 5.0, 5.0]], np.float32), np.zeros(1, np.uint8))
    bad = ParticleCollection.concatenate([c.copy(), far])
    srv = Server(gpu_context)
    ids = [srv.submit(dd, cc.copy(), s) for dd, cc, s in jobs[:1]]
    bad_id = srv.submit(d, bad, 4)
    ids += [srv.submit(dd, cc.copy(), s) for dd, cc, s in jobs[1:]]
    print("case serving key range", flush=True)
    assert srv.tick() == 4
    info = srv.job(bad_id)
    assert info.state == abi.JOB_FAILED and info.status == abi.FSA_E_INVALID
    while srv.stats()["jobs_running"] or srv.stats()["jobs_queued"]:
        srv.tick()
    for jid, (dd, cc, steps) in zip(ids, jobs):
        assert srv.job(jid).state == abi.JOB_DONE
        np.testing.assert_array_equal(srv.collect(jid).movement, _alone(gpu_context, dd, cc, steps).movement)
    assert srv.stats()["jobs_failed"] == 1
    srv.release(bad_id)
    with pytest.raises(Exception):
        srv.job(bad_id)
    srv.close()
