"""Solver requests and results in the reference's wire formats.

A request is the JSON object doSolvingFFT receives (do_solving_fft.jl:1-158):
``mesherOutput`` (``mesher_matrices[name][x][y][z]`` booleans per material,
``cell_size``, ``origin``, ``n_cells``), ``solverInput`` (``materials``,
``ports``, ``lumped_elements``, ``frequencies``, ``unit``,
``ports_scattering_value``), ``solverAlgoParams`` (``innerIteration``,
``outerIteration``, ``convergenceThreshold``) and ``solverType`` (1
quasi-static, 2 retarded).  The result is dump_json_data's dictionary
(format_input_output_solver_functions.jl:1-31): per port pair, the
frequency series of Z, S and Y as [re, im] pairs, JSON-encoded.

Conventions kept from the reference:
  * lengths are scaled by the unit (getEscalFrom, :190-210), the origin is
    not (do_solving_fft.jl:9-11);
  * a material's loss tangent is its ``tangent_delta_conductivity``
    (material(), :38-64; mesher_FFT.jl:21 makes epsr = eps_re (1 - j tan_D));
  * a grid is a conductor when its material's conductivity is positive
    (is_material_conductor.jl);
  * lumped elements take R, L and C from ``rlcParams`` (read_lumped_elements,
    :124-188).
Materials pair with grids by name; the reference pairs them by position
(its material list against the Dict order of ``mesher_matrices``).
"""
import gzip
import json

import numpy as np

from . import mesher, voxel_mesher

UNITS = {"m": 1.0, "dm": 1e-1, "cm": 1e-2, "mm": 1e-3, "microm": 1e-6, "nanom": 1e-9}


class Request:
    """A parsed request: the mesher's and the solver's inputs."""


def load(path_or_obj):
    """The request object of a JSON (optionally gzipped) file, or the object itself."""
    if isinstance(path_or_obj, dict):
        return path_or_obj
    opener = gzip.open if str(path_or_obj).endswith(".gz") else open
    with opener(path_or_obj, "rt") as f:
        return json.load(f)


def escal_of(unit):
    """getEscalFrom (format_input_output_solver_functions.jl:190-210): an
    unknown unit scales by 1."""
    return UNITS.get(unit, 1.0)


def material(el):
    """material() (format_input_output_solver_functions.jl:38-64) as a
    mesher.Material: sigmar = conductivity, eps_re = permittivity, tan_D =
    tangent_delta_conductivity."""
    return mesher.Material(el["name"], sigma=el["conductivity"], eps_re=el["permittivity"],
                           tan_d=el.get("tangent_delta_conductivity", 0))


def parse(obj):
    """Parse a request object (do_solving_fft.jl:1-64, read_ports,
    read_lumped_elements)."""
    d = load(obj)
    mo, si = d["mesherOutput"], d["solverInput"]
    R = Request()
    R.escal = escal_of(si["unit"])
    cs = mo["cell_size"]
    R.cell = (cs["cell_size_x"] * R.escal, cs["cell_size_y"] * R.escal, cs["cell_size_z"] * R.escal)
    og = mo["origin"]
    R.origin = (float(og["origin_x"]), float(og["origin_y"]), float(og["origin_z"]))
    by_name = {m["name"]: m for m in si["materials"]}
    R.names, R.grids, R.materials = [], [], []
    for name, value in mo["mesher_matrices"].items():
        if name not in by_name:
            raise ValueError("mesher_matrices names material %r, which solverInput.materials does not list" % name)
        R.names.append(name)
        R.grids.append(np.asarray(value, dtype=bool))
        R.materials.append(material(by_name[name]))
    shapes = {g.shape for g in R.grids}
    if len(shapes) != 1:
        raise ValueError("mesher_matrices grids differ in shape: %s" % sorted(shapes))
    nc = mo.get("n_cells")
    if nc is not None and tuple(R.grids[0].shape) != (nc["n_cells_x"], nc["n_cells_y"], nc["n_cells_z"]):
        raise ValueError("mesher_matrices shape %s is not n_cells %s" % (R.grids[0].shape, nc))
    R.dominant = [m.sigmar > 0 for m in R.materials]
    R.freqs = np.asarray(si["frequencies"], dtype=np.float64)
    R.ports = [(np.asarray(p["inputElement"], dtype=np.float64) * R.escal,
                np.asarray(p["outputElement"], dtype=np.float64) * R.escal) for p in si["ports"]]
    R.port_names = [p.get("name", "") for p in si["ports"]]
    R.lumped_elements = []
    for e in si.get("lumped_elements", []):
        rlc = e.get("rlcParams", {}) or {}
        R.lumped_elements.append(dict(start=np.asarray(e["inputElement"], dtype=np.float64) * R.escal,
                                      end=np.asarray(e["outputElement"], dtype=np.float64) * R.escal,
                                      type=int(e["type"]), value=float(e.get("value", 0.0)),
                                      R=float(rlc.get("resistance", 0.0)), L=float(rlc.get("inductance", 0.0)),
                                      C=float(rlc.get("capacitance", 0.0))))
    R.R0 = float(si.get("ports_scattering_value", 50.0))
    ap = d.get("solverAlgoParams", {})
    R.inner_iter = int(ap.get("innerIteration", 100))
    R.outer_iter = int(ap.get("outerIteration", 1))
    R.tol = float(ap.get("convergenceThreshold", 1e-4))
    R.solver_type = int(d.get("solverType", 1))
    return R


def mesh_of(R, p_self_fn=None):
    """The reference mesher (mesher_FFT.jl) on a parsed request."""
    return voxel_mesher.mesh_request(R.grids, R.materials, R.cell, R.origin, R.dominant, R.ports, R.lumped_elements,
                                     p_self_fn=p_self_fn)


def dump_json_data(Z, S, Y, num_ports, id_, partial=False, freq_index=None):
    """dump_json_data (format_input_output_solver_functions.jl:1-31): for each
    port pair (i, then j), [[ [re, im] per frequency ]], JSON-encoded."""
    def enc(A):
        A = np.asarray(A, dtype=np.complex128)
        out = []
        for i in range(num_ports):
            for j in range(num_ports):
                out.append([[[float(v.real), float(v.imag)] for v in A[i, j, :]]])
        return json.dumps(out)

    return {"matrices": {"matrix_Z": enc(Z), "matrix_S": enc(S), "matrix_Y": enc(Y)}, "isStopped": False, "id": id_,
            "partial": partial, "freqIndex": freq_index}


def parse_matrix(s, num_ports, nfreq):
    """Inverse of one dump_json_data matrix: (P, P, F) complex."""
    data = json.loads(s)
    A = np.zeros((num_ports, num_ports, nfreq), dtype=np.complex128)
    for idx, entry in enumerate(data):
        i, j = divmod(idx, num_ports)
        A[i, j, :] = [complex(a, b) for a, b in entry[0]]
    return A


def solve(obj, id_="request", ctx=None, p_self_fn=None, solver_kwargs=None, record=None):
    """doSolvingFFT on the device: parse, mesh, kernel spectra, the
    frequency sweep of FFT_solver_QS_S_type (solverType 1 quasi-static, 2
    retarded), and dump_json_data's result (None when stopped).  record:
    a list that receives each solve's flag and iteration counts."""
    from . import peec, sweep

    R = parse(obj)
    ctx = ctx or peec.Context(0)
    if p_self_fn is None:
        def p_self_fn(sx, sy, sz):
            c = np.zeros((1, 3))
            return [float(np.ravel(peec.compute_row_P_sup(c[0], c, sx, sy, sz, cs, cs, ctx=ctx))[0]) for cs in (1, 3, 5)]
    M = mesh_of(R, p_self_fn=p_self_fn)
    kw = dict(solver_kwargs or {})
    if R.solver_type == 2:
        kw.setdefault("QS_Rcc_FW", 2)
    S = sweep.Solver(M, ctx=ctx, **kw)
    out = sweep.FFT_solver_QS_S_type(R.freqs, M, R.inner_iter, R.tol, ports_scatter_value=R.R0, solver=S,
                                     record=record)
    if out is None:
        return None
    return dump_json_data(out["Z"], out["S"], out["Y"], len(R.ports), id_)
