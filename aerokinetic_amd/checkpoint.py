"""Device checkpoints of a uniform or two-level case.

Restates the reference's restart checkpoints (src/aerokinetic/checkpoint.py:32-135,
refined_checkpoint.py:54-142) for the HIP state.  The reference stores lbmpy field
arrays and boundary index lists, which only lbmpy can read; here a checkpoint holds
the library's restartable device state (`ak_save_state`: the raw population buffer in
its storage precision, the AA parity and the boundary histories: characteristic
incoming/derivative/clock/glancing trace, legacy outflow pdf/pdf_nd), plus the same
metadata: project, case fingerprint, step count, wall fallbacks and the run context.
Saving only macroscopic fields or exported populations would not reproduce the
history-dependent boundaries, so a restored case continues bit-identically.

Format: an uncompressed .npz with `state` (uint8) and `metadata` (JSON), replaced
atomically.
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

from . import _lib
from ._io import PublishedFile

FORMAT = "aerokinetic_mi355x_checkpoint"
SCHEMA = 1


def _library():
    return _lib.load().ak_version().decode()


def save_checkpoint(solver, path, context=None):
    """Write the restartable state of `solver` (Solver or RefinedSolver) to `path`."""
    from .project import Project
    from .refined import RefinedSolver
    from .refined_plan import refined_project_dict
    L = _lib.load()
    refined = isinstance(solver, RefinedSolver)
    if refined:
        n = int(L.ak_refined_state_bytes(solver.h))
        state = np.empty(n, dtype=np.uint8)
        _lib.check(L.ak_refined_save_state(solver.h, _lib._vp(state.ctypes.data), n), None, "ak_refined_save_state")
        project = refined_project_dict(solver.project, solver.refinement)
        fingerprint = solver.case_fingerprint
    else:
        h = solver.lattice.h
        n = int(L.ak_state_bytes(h))
        state = np.empty(n, dtype=np.uint8)
        _lib.check(L.ak_save_state(h, _lib._vp(state.ctypes.data), n), h, "ak_save_state")
        p = Project(solver.section, solver.flow, solver.grid)
        project, fingerprint = p.as_dict(), p.case_fingerprint
    meta = {"format": FORMAT, "schema_version": SCHEMA, "kind": "refined" if refined else "uniform",
            "library": _library(), "storage": solver.storage, "project": project, "case_fingerprint": fingerprint,
            "steps": int(solver.steps), "wall_fallback_links": int(solver.wall_fallback_links),
            "context": context or {}}
    # the plan, so that a case built on an explicit lattice is rebuilt on the same one
    meta["plan"] = json.loads(json.dumps(solver.plan.as_dict(), default=list))
    if not refined:
        meta["plan_shape"] = list(solver.plan.shape)
    arrays = {"state": state, "metadata": np.asarray(json.dumps(meta, ensure_ascii=False, allow_nan=False))}
    with PublishedFile(path) as handle:
        np.savez(handle, **arrays)
    return Path(path)


def read_metadata(path):
    with np.load(path, allow_pickle=False) as saved:
        meta = json.loads(str(saved["metadata"]))
    if meta.get("format") != FORMAT or meta.get("schema_version") != SCHEMA:
        raise ValueError("Unsupported checkpoint format")
    return meta


def restore_checkpoint(solver_type, path, progress=None, device=0, plan=None):
    """Rebuild the case of a checkpoint on the plan it recorded (or `plan`) and load its
    device state (same storage)."""
    from .config import Plan
    from .project import Project
    from .refined import RefinedSolver
    from .refined_plan import RefinedPlan, RefinementSettings
    with np.load(path, allow_pickle=False) as saved:
        meta = json.loads(str(saved["metadata"]))
        state = np.ascontiguousarray(saved["state"])
    if meta.get("format") != FORMAT or meta.get("schema_version") != SCHEMA:
        raise ValueError("Unsupported checkpoint format")
    if meta.get("library") != _library():
        raise ValueError("Checkpoint written by another library build; use that build to resume")
    if type(meta.get("steps")) is not int or meta["steps"] < 0:
        raise ValueError("Invalid checkpoint step count")
    L = _lib.load()
    if issubclass(solver_type, RefinedSolver):
        if meta["kind"] != "refined":
            raise ValueError("Not a two-level checkpoint")
        project = Project.from_dict(meta["project"]["project"])
        settings = RefinementSettings(**meta["project"]["refinement"])
        if plan is None:
            plan = RefinedPlan.from_dict(meta["plan"]) if "plan" in meta else RefinedPlan.create(project, settings)
        solver = solver_type(project, settings, progress, storage=meta["storage"], device=device, plan=plan)
        if solver.case_fingerprint != meta["case_fingerprint"]:
            raise ValueError("Checkpoint physical/numerical case is inconsistent")
        _lib.check(L.ak_refined_load_state(solver.h, _lib._vp(state.ctypes.data), state.size), None,
                   "ak_refined_load_state")
    else:
        if meta["kind"] != "uniform":
            raise ValueError("Not a uniform checkpoint")
        project = Project.from_dict(meta["project"])
        if project.case_fingerprint != meta["case_fingerprint"]:
            raise ValueError("Checkpoint physical/numerical case is inconsistent")
        if plan is None and "plan" in meta:
            plan = Plan.from_dict(meta["plan"])
        solver = solver_type(project.section, project.flow, project.grid, progress, storage=meta["storage"],
                             device=device, plan=plan)
        if list(solver.plan.shape) != meta["plan_shape"]:
            raise ValueError("Checkpoint lattice shape differs from the rebuilt case")
        h = solver.lattice.h
        _lib.check(L.ak_load_state(h, _lib._vp(state.ctypes.data), state.size), h, "ak_load_state")
    if solver.wall_fallback_links != meta["wall_fallback_links"]:
        raise ValueError("Checkpoint wall links differ from the rebuilt case")
    solver.steps = meta["steps"]
    solver._checkpoint_context = meta["context"]
    solver._restored_control = meta["context"].get("run_control")
    solver.diagnostics()
    return solver
