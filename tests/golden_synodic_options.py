"""Detection options of the synodic golden vectors (shared with
tests/golden/make_synodic_golden.py, which imports this list)."""

OPTIONS = [
    dict(normal=[0, 1, 0, 0, 0, 0], interp_kind="linear", segment_refine=0),
    dict(normal=[0, 1, 0, 0, 0, 0], interp_kind="cubic", segment_refine=0, newton_max_iter=4),
    dict(normal=[0, 1, 0, 0, 0, 0], interp_kind="cubic", segment_refine=0, newton_max_iter=25),
    dict(normal=[0, 1, 0, 0, 0, 0], interp_kind="cubic", segment_refine=20, newton_max_iter=25),
    dict(normal=[0, 1, 0, 0, 0, 0], interp_kind="linear", segment_refine=20),
    dict(normal=[0, 1, 0, 0, 0, 0], interp_kind="cubic", segment_refine=20, newton_max_iter=25,
         direction=1),
    dict(normal=[0, 1, 0, 0, 0, 0], interp_kind="cubic", segment_refine=0, newton_max_iter=25,
         direction=-1),
    dict(normal=[1, 0, 0, 0, 0, 0], offset=0.85, interp_kind="cubic", segment_refine=20,
         newton_max_iter=25, plane_coords=("y", "vy")),
    dict(normal=[0, 1, 0, 0, 0, 0], interp_kind="cubic", segment_refine=5, newton_max_iter=25,
         max_hits_per_traj=2, dedup_point_tol=1e-3),
]
