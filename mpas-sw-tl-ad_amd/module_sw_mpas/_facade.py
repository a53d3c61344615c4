"""The ``mpas_sw_module`` object: the reference's Fortran module interface
(``src/mpas_sw_module.f90``) as f2py exposes it, with the stepping routines
on the GPU.

Attribute semantics follow f2py (numpy/f2py/src/fortranobject.c):

* names are lower case; module parameters (``a``, ``omega``, ``gravity``,
  ...) are readable;
* reading an allocated array returns a Fortran-ordered numpy array that
  aliases the module's storage, so ``mpsw.u[0, 0] = x`` changes the state;
  reading an unallocated one returns None;
* assigning an array (re)allocates it with the assigned shape and copies,
  cast to the variable's type;
* assigning an unknown name is accepted and stored aside (the reference
  driver relies on this for ``config_calendar_type``,
  ``config_h_ScaleWithMesh`` and the ``*solve*`` ranges).

The host arrays are the source of truth between calls: every ``sw_rk4*``
uploads the level-1 state, steps on the GPU in the library's internal order,
and writes both time levels back.  Diagnostic arrays are recomputed on the
GPU when they are read after a step.
"""

from __future__ import annotations

import numpy as np

# ---- module parameters (src/mpas_sw_module.f90:4-18)
_PARAMS = {
    "rkind": 8, "strkind": 512,
    "pii": 3.141592653589793, "a": 6371229.0, "omega": 7.29212e-5, "gravity": 9.80616,
    "rgas": 287.0, "rv": 461.6,
}
_PARAMS["rvord"] = _PARAMS["rv"] / _PARAMS["rgas"]
_PARAMS["cp"] = 7.0 * _PARAMS["rgas"] / 2.0
_PARAMS["cv"] = _PARAMS["cp"] - _PARAMS["rgas"]
_PARAMS["cvpm"] = -_PARAMS["cv"] / _PARAMS["cp"]
_PARAMS["prandtl"] = 1.0

# ---- scalar variables with their initial values (:23-55, :73-75, :117, :206-207)
_SCALARS = {
    "config_test_case": 5, "config_time_integration": "RK4", "config_dt": 150.0,
    "config_start_time": "2000-01-01_00:00:00", "config_stop_time": "none",
    "config_run_duration": "none", "config_stats_interval": 100,
    "config_h_scalewithmesh": False, "config_h_mom_eddy_visc2": 0.0,
    "config_h_mom_eddy_visc4": 0.0, "config_h_tracer_eddy_diff2": 0.0,
    "config_h_tracer_eddy_diff4": 0.0, "config_thickness_adv_order": 2,
    "config_tracer_adv_order": 2, "config_positive_definite": False, "config_monotonic": False,
    "config_wind_stress": False, "config_bottom_drag": False, "config_apvm_upwinding": 0.5,
    "config_num_halos": 2,
    "nvertlevels": 1, "ncells": 0, "nedges": 0, "nvertices": 0, "ntracers": 0,
    "cellstart": 0, "cellend": 0, "vertexstart": 0, "vertexend": 0, "edgestart": 0,
    "edgeend": 0, "vertexdegree": 0, "ncellssolve": 0, "nedgessolve": 0, "maxedges": 0,
    "maxedges2": 0, "cellsolvestart": 0, "cellsolveend": 0, "edgesolvestart": 0,
    "edgesolveend": 0, "r3": 0,
    "on_a_sphere": False, "is_periodic": False, "x_period": 0.0, "y_period": 0.0,
    "sphere_radius": 0.0, "rk_step": 0, "lhs": 0.0, "rhs": 0.0, "pt": 0.0,
}

_INT_ARRAYS = {"boundarycell", "boundaryedge", "advcells", "cellsoncell", "cellsonedge",
               "cellsonvertex", "edgesoncell", "edgesonedge", "edgesonvertex", "verticesoncell",
               "verticesonedge", "nedgesoncell", "nedgesonedge"}

_DIAG = ("v", "h_edge", "h_vertex", "pv_edge", "pv_vertex", "pv_cell", "circulation", "vorticity",
         "divergence", "ke", "vh", "vorticity_cell", "gradpvn", "gradpvt")
_RECON = ("ureconstructx", "ureconstructy", "ureconstructz", "ureconstructzonal",
          "ureconstructmeridional")
_GPU_DIAG_NAME = {"gradpvn": "gradPVn", "gradpvt": "gradPVt"}


class _Module:
    """Emulates the f2py Fortran-module object ``mpas_sw_module``."""

    def __init__(self):
        object.__setattr__(self, "_vars", dict(_SCALARS))
        object.__setattr__(self, "_arrays", {})
        object.__setattr__(self, "_extra", {})
        object.__setattr__(self, "_gpu", None)
        object.__setattr__(self, "_diag_state", None)   # (tag, h, u) the diagnostics describe
        object.__setattr__(self, "_diag_fresh", set())

    # ------------------------------------------------------------ attributes
    def __getattr__(self, name):
        key = name.lower()
        d = object.__getattribute__(self, "__dict__")
        if key in _PARAMS:
            return _PARAMS[key]
        if key in d["_arrays"]:
            self._refresh_derived(key)
            return d["_arrays"][key]
        if key in d["_vars"]:
            return d["_vars"][key]
        if key in _ALL_ARRAYS:
            return None                      # declared but not allocated
        if name in d["_extra"]:
            return d["_extra"][name]
        raise AttributeError(name)

    def __setattr__(self, name, value):
        key = name.lower()
        if name != key:
            # f2py exposes lower-case names only; a mixed-case assignment
            # lands in the object's dict and never reaches the module
            self._extra[name] = value
            return
        if key in _PARAMS:
            raise AttributeError(f"{name} is a module parameter")
        if key in _ALL_ARRAYS:
            dt = np.int32 if key in _INT_ARRAYS else np.float64
            self._arrays[key] = np.array(value, dtype=dt, order="F", copy=True)
            self._diag_fresh.discard(key)
            return
        if key in self._vars:
            old = self._vars[key]
            if isinstance(old, bool):
                value = bool(value)
            elif isinstance(old, int):
                value = int(value)
            elif isinstance(old, float):
                value = float(value)
            elif isinstance(old, str):
                value = str(value)
            self._vars[key] = value
            return
        self._extra[name] = value

    def __dir__(self):
        return sorted(set(_PARAMS) | set(self._vars) | set(_ALL_ARRAYS) | set(_ROUTINES))

    # -------------------------------------------------------------- helpers
    def _dims(self):
        v = self._vars
        return (v["nvertlevels"], v["ncells"], v["nedges"], v["nvertices"], v["ntracers"],
                v["maxedges"], v["maxedges2"], v["vertexdegree"], v["r3"])

    def _alloc(self, name, shape, dtype=np.float64):
        if name in self._arrays:
            raise RuntimeError(f"Fortran runtime error: attempting to allocate already allocated variable '{name}'")
        self._arrays[name] = np.zeros(shape, dtype=dtype, order="F")

    # --------------------------------------------------------- allocation
    def var_allocation(self):
        """src/mpas_sw_module.f90:211-367"""
        nvl, C, E, V, nt, mE, mE2, vd, r3 = self._dims()
        al = self._alloc
        al("u", (2, nvl, E)); al("h", (2, nvl, C)); al("tracers", (2, nt, nvl, C))
        al("u_src", (nvl, E))
        for n in ("h_s", "fcell", "areacell", "meshdensity", "latcell", "loncell", "xcell", "ycell", "zcell"):
            al(n, (C,))
        for n in ("fvertex", "areatriangle", "latvertex", "lonvertex", "xvertex", "yvertex", "zvertex"):
            al(n, (V,))
        for n in ("fedge", "dvedge", "dcedge", "meshscalingdel2", "meshscalingdel4", "latedge",
                  "lonedge", "xedge", "yedge", "zedge", "angleedge"):
            al(n, (E,))
        al("weightsonedge", (mE2, E)); al("kiteareasonvertex", (vd, V))
        al("boundarycell", (nvl, C), np.int32); al("boundaryedge", (nvl, E), np.int32)
        al("advcells", (21, C), np.int32); al("cellsoncell", (mE, C), np.int32)
        al("cellsonedge", (2, E), np.int32); al("cellsonvertex", (vd, V), np.int32)
        al("edgesoncell", (mE, C), np.int32); al("edgesonedge", (mE2, E), np.int32)
        al("edgesonvertex", (vd, V), np.int32); al("verticesoncell", (mE, C), np.int32)
        al("verticesonedge", (2, E), np.int32); al("nedgesoncell", (C,), np.int32)
        al("nedgesonedge", (E,), np.int32)
        al("deriv_two", (15, 2, E))
        al("defc_a", (mE, C)); al("defc_b", (mE, C))
        al("localverticalunitvectors", (r3, C)); al("edgenormalvectors", (r3, E))
        al("celltangentplane", (r3, 2, C)); al("coeffs_reconstruct", (r3, mE, C))
        self._alloc_diag("", nvl, C, E, V, nt)

    def _alloc_diag(self, sfx, nvl, C, E, V, nt):
        for n in _RECON:
            self._alloc(n + sfx, (nvl, C))
        kinds = {"v": E, "h_edge": E, "h_vertex": V, "pv_edge": E, "pv_vertex": V, "pv_cell": C,
                 "circulation": V, "vorticity": V, "divergence": C, "ke": C, "vh": E,
                 "vorticity_cell": C, "gradpvn": E, "gradpvt": E}
        for n, size in kinds.items():
            self._alloc(n + sfx, (nvl, size))
        self._alloc("tend_u" + sfx, (nvl, E))
        self._alloc("tend_h" + sfx, (nvl, C))
        self._alloc("tend_tracer" + sfx, (nt, nvl, C))

    def var_allocation_tlm(self):
        """src/mpas_sw_module.f90:369-429"""
        nvl, C, E, V, nt, *_ = self._dims()
        self._alloc("u_tl", (2, nvl, E)); self._alloc("h_tl", (2, nvl, C))
        self._alloc("tracers_tl", (2, nt, nvl, C))
        self._alloc_diag("_tl", nvl, C, E, V, nt)

    def var_allocation_adj(self):
        """src/mpas_sw_module.f90:431-493"""
        nvl, C, E, V, nt, *_ = self._dims()
        self._alloc("u_ad", (2, nvl, E)); self._alloc("h_ad", (2, nvl, C))
        self._alloc("tracers_ad", (2, nt, nvl, C))
        self._alloc_diag("_ad", nvl, C, E, V, nt)

    # --------------------------------------------------------- test cases
    def _mesh(self):
        from mpas_sw_amd.mesh import Mesh
        A = self._arrays
        g = lambda n: np.ascontiguousarray(A[n])
        return Mesh(
            level=-1, nCells=self._vars["ncells"], nEdges=self._vars["nedges"],
            nVertices=self._vars["nvertices"],
            xCell=g("xcell"), yCell=g("ycell"), zCell=g("zcell"), latCell=g("latcell"), lonCell=g("loncell"),
            xEdge=g("xedge"), yEdge=g("yedge"), zEdge=g("zedge"), latEdge=g("latedge"), lonEdge=g("lonedge"),
            xVertex=g("xvertex"), yVertex=g("yvertex"), zVertex=g("zvertex"), latVertex=g("latvertex"),
            lonVertex=g("lonvertex"), cellsOnEdge=g("cellsonedge"), verticesOnEdge=g("verticesonedge"),
            nEdgesOnCell=g("nedgesoncell"), edgesOnCell=g("edgesoncell"), verticesOnCell=g("verticesoncell"),
            cellsOnCell=g("cellsoncell"), cellsOnVertex=g("cellsonvertex"), edgesOnVertex=g("edgesonvertex"),
            kiteAreasOnVertex=g("kiteareasonvertex"), nEdgesOnEdge=g("nedgesonedge"),
            edgesOnEdge=g("edgesonedge"), weightsOnEdge=g("weightsonedge"), dcEdge=g("dcedge"),
            dvEdge=g("dvedge"), angleEdge=g("angleedge"), areaCell=g("areacell"),
            areaTriangle=g("areatriangle"), meshDensity=g("meshdensity"),
            maxEdges=self._vars["maxedges"], maxEdges2=self._vars["maxedges2"],
            vertexDegree=self._vars["vertexdegree"])

    def _test_case(self, case):
        from mpas_sw_amd.testcases import initial_conditions
        ic = initial_conditions(self._mesh(), case)
        A = self._arrays
        for src, dst in (("xCell", "xcell"), ("yCell", "ycell"), ("zCell", "zcell"),
                         ("xVertex", "xvertex"), ("yVertex", "yvertex"), ("zVertex", "zvertex"),
                         ("xEdge", "xedge"), ("yEdge", "yedge"), ("zEdge", "zedge"),
                         ("dvEdge", "dvedge"), ("dcEdge", "dcedge"), ("areaCell", "areacell"),
                         ("areaTriangle", "areatriangle"), ("kiteAreasOnVertex", "kiteareasonvertex")):
            A[dst][...] = ic[src]
        A["u"][0, 0, :] = ic["u"]
        A["h"][0, 0, :] = ic["h"]
        if case in (2, 5):
            A["fedge"][...] = ic["fEdge"]
            A["fvertex"][...] = ic["fVertex"]
        if case == 5:
            A["h_s"][...] = ic["h_s"]
            A["loncell"][...] = ic["lonCell"]
            A["tracers"][0, 0, 0, :] = ic["tracer"]
        self._gpu_reset()

    def sw_test_case_1(self):
        """src/mpas_sw_testcases.f90:1-66"""
        self._test_case(1)

    def sw_test_case_2(self):
        """src/mpas_sw_testcases.f90:68-145"""
        self._test_case(2)

    def sw_test_case_5(self):
        """src/mpas_sw_testcases.f90:147-254"""
        self._test_case(5)

    def sw_test_case_6(self):
        """src/mpas_sw_testcases.f90:256-319"""
        self._test_case(6)

    @staticmethod
    def sphere_distance(lat1, lon1, lat2, lon2, radius):
        from mpas_sw_amd.testcases import sphere_distance
        return sphere_distance(lat1, lon1, lat2, lon2, radius)

    @staticmethod
    def aa(theta):
        from mpas_sw_amd.testcases import aa
        return aa(theta)

    @staticmethod
    def bb(theta):
        from mpas_sw_amd.testcases import bb
        return bb(theta)

    @staticmethod
    def cc(theta):
        from mpas_sw_amd.testcases import cc
        return cc(theta)

    # ---------------------------------------------------------- GPU binding
    def _gpu_reset(self):
        object.__setattr__(self, "_gpu", None)

    def _model(self):
        """The GPU model for the current mesh and configuration."""
        v = self._vars
        A = self._arrays
        if v["config_thickness_adv_order"] in (3, 4) and np.any(A["deriv_two"] != 0):
            raise NotImplementedError("thickness advection orders 3/4 with a non-zero deriv_two")
        key = (v["config_dt"], v["config_apvm_upwinding"], v["config_test_case"],
               v["config_h_mom_eddy_visc2"], v["config_h_mom_eddy_visc4"],
               bool(v["config_wind_stress"]), bool(v["config_bottom_drag"]))
        gpu = self._gpu
        if gpu is not None and gpu["key"][2] != key[2]:
            gpu = None      # another test case: other h_s/fVertex, frozen u
        if gpu is None:
            from mpas_sw_amd.core import SWModel
            A = self._arrays
            phys = {"dcEdge": A["dcedge"], "dvEdge": A["dvedge"], "areaCell": A["areacell"],
                    "areaTriangle": A["areatriangle"], "kiteAreasOnVertex": A["kiteareasonvertex"],
                    "fVertex": A["fvertex"], "h_s": A["h_s"], "lonCell": A["loncell"],
                    "coeffs_reconstruct": A["coeffs_reconstruct"]}
            m = SWModel(self._mesh(), phys, v["config_dt"], apvm=v["config_apvm_upwinding"],
                        test_case=v["config_test_case"], visc2=v["config_h_mom_eddy_visc2"],
                        visc4=v["config_h_mom_eddy_visc4"],
                        wind_stress=bool(v["config_wind_stress"]),
                        bottom_drag=bool(v["config_bottom_drag"]))
            # the last RK stage's tendencies, which the module arrays tend_h,
            # tend_u, tend_h_tl, tend_u_tl hold after each step
            # (src/mpas_sw_module.f90:188-195, :577-586, :700-712, :838-848)
            import torch
            tend = tuple(torch.zeros(n, dtype=torch.float64, device=m.device)
                         for n in (m.nC, m.nE, m.nC, m.nE))
            m.set_tend_output(*tend)
            gpu = {"model": m, "key": key, "tend": tend}
            object.__setattr__(self, "_gpu", gpu)
        elif gpu["key"] != key:
            gpu["model"].set_config(dt=v["config_dt"], apvm=v["config_apvm_upwinding"],
                                    visc2=v["config_h_mom_eddy_visc2"],
                                    visc4=v["config_h_mom_eddy_visc4"],
                                    wind_stress=bool(v["config_wind_stress"]),
                                    bottom_drag=bool(v["config_bottom_drag"]))
            gpu["key"] = key
        if v["config_wind_stress"]:
            # the reference reads the module array u_src at every tendency call
            gpu["model"].set_forcing(A["u_src"][0])
        # meshScalingDel2/4 scale the del2/del4 terms per edge; they are 1
        # unless config_h_ScaleWithMesh was set for compute_mesh_scaling
        # (src/mpas_sw_dyncore.f90:362-392); uploaded when they change
        d2, d4 = A["meshscalingdel2"], A["meshscalingdel4"]
        scaled = bool(np.any(d2 != 1.0) or np.any(d4 != 1.0))
        sig = (d2.tobytes(), d4.tobytes()) if scaled else None
        if gpu.get("scaling") != sig:
            if scaled:
                gpu["model"].set_mesh_scaling(d2, d4)
            else:
                gpu["model"].set_mesh_scaling()
            gpu["scaling"] = sig
        return gpu["model"]

    def _up(self, m, arr, kind):
        return m.to_internal(np.asarray(arr), kind, member=False)

    def _tend_back(self, m, names):
        """Copy the tendencies the last step left on the GPU into the module
        arrays `names` (those that are allocated)."""
        th, tu, tht, tut = self._gpu["tend"]
        src = {"tend_h": (th, "c"), "tend_u": (tu, "e"), "tend_h_tl": (tht, "c"), "tend_u_tl": (tut, "e")}
        for n in names:
            if n in self._arrays:
                t, kind = src[n]
                self._arrays[n][0, :] = m.to_reference(t, kind)

    def _set_diag_state(self, tag, h, u, u_tl=None):
        object.__setattr__(self, "_diag_state", (tag, h, u, u_tl))
        self._diag_fresh.clear()

    def _refresh_derived(self, key):
        """Fill a diagnostic/reconstruction array from the GPU state it describes."""
        if key in self._diag_fresh or self._diag_state is None or self._gpu is None:
            return
        base = key[:-3] if key.endswith(("_tl", "_ad")) else key
        sfx = key[len(base):]
        if base not in _DIAG and base not in _RECON:
            return
        tag, h, u, u_tl = self._diag_state
        m = self._gpu["model"]
        A = self._arrays
        if sfx == "":
            d = m.diagnostics(h, u)
            for n in _DIAG:
                nm = _GPU_DIAG_NAME.get(n, n)
                if nm in d and n in A:
                    t, kind = d[nm]
                    A[n][0, :] = m.to_reference(t, kind)
                    self._diag_fresh.add(n)
            r = m.reconstruct(u.to(m.dtype) if u.dtype != m.dtype else u)
            for i, n in enumerate(_RECON):
                A[n][0, :] = m.to_reference(r[i], "c")
                self._diag_fresh.add(n)
        elif sfx == "_tl" and u_tl is not None and base in _RECON:
            r = m.reconstruct(u_tl)
            for i, n in enumerate(_RECON):
                if n + "_tl" in A:
                    A[n + "_tl"][0, :] = m.to_reference(r[i], "c")
                    self._diag_fresh.add(n + "_tl")
        self._diag_fresh.add(key)

    # ------------------------------------------------------------ routines
    def sw_mpas_init_block(self):
        """src/mpas_sw_module.f90:495-526: initial diagnostics, mesh scaling,
        unit vectors, reconstruction coefficients and reconstruction."""
        from mpas_sw_amd.reconstruct import initialize_vectors, init_reconstruct
        A = self._arrays
        A["meshscalingdel2"][...] = 1.0
        A["meshscalingdel4"][...] = 1.0
        if self._vars["config_h_scalewithmesh"]:
            c1 = A["cellsonedge"][0] - 1
            c2 = A["cellsonedge"][1] - 1
            # compute_mesh_scaling (src/mpas_sw_dyncore.f90:362-392), pow of the
            # C library element by element, as the reference's compiled code calls it
            import math
            avg = ((A["meshdensity"][c1] + A["meshdensity"][c2]) / 2.0).tolist()
            A["meshscalingdel2"][...] = [1.0 / math.pow(x, 5.0 / 12.0) for x in avg]
            A["meshscalingdel4"][...] = [1.0 / math.pow(x, 5.0 / 6.0) for x in avg]
        xc = np.stack([A["xcell"], A["ycell"], A["zcell"]], 1)
        xe = np.stack([A["xedge"], A["yedge"], A["zedge"]], 1)
        lvu, env, ctp = initialize_vectors(xc, xe, A["cellsonedge"], A["edgesoncell"])
        A["localverticalunitvectors"][...] = lvu.T
        A["edgenormalvectors"][...] = env.T
        A["celltangentplane"][...] = np.transpose(ctp, (2, 1, 0))
        A["coeffs_reconstruct"][...] = init_reconstruct(xc, xe, env, ctp, A["nedgesoncell"], A["edgesoncell"])
        self._gpu_reset()
        m = self._model()
        self._set_diag_state("init", self._up(m, A["h"][0, 0], "c"), self._up(m, A["u"][0, 0], "e"))

    def sw_rk4(self):
        """src/mpas_sw_module.f90:528-635 (one step on the GPU)."""
        m = self._model()
        A = self._arrays
        h = self._up(m, A["h"][0, 0], "c")
        u = self._up(m, A["u"][0, 0], "e")
        m.base(h, u, 1)
        hn, un = m.to_reference(h, "c"), m.to_reference(u, "e")
        A["h"][0, 0, :] = hn; A["h"][1, 0, :] = hn
        A["u"][0, 0, :] = un; A["u"][1, 0, :] = un
        A["tracers"][1] = A["tracers"][0]
        self._tend_back(m, ("tend_h", "tend_u"))
        self._set_diag_state("fwd", h, u)

    def sw_rk4_tlm(self):
        """src/mpas_sw_module.f90:637-781 (one step on the GPU; advances the
        nonlinear state too, :774-777)."""
        m = self._model()
        A = self._arrays
        if "u_tl" not in A:
            raise RuntimeError("var_allocation_tlm() has not been called")
        h = self._up(m, A["h"][0, 0], "c")
        u = self._up(m, A["u"][0, 0], "e")
        ht = self._up(m, A["h_tl"][0, 0], "c")
        ut = self._up(m, A["u_tl"][0, 0], "e")
        m.tlm(h, u, ht, ut, 1)
        for name, t, kind in (("h", h, "c"), ("u", u, "e"), ("h_tl", ht, "c"), ("u_tl", ut, "e")):
            x = m.to_reference(t, kind)
            A[name][0, 0, :] = x
            A[name][1, 0, :] = x
        A["tracers"][1] = A["tracers"][0]
        A["tracers_tl"][1] = A["tracers_tl"][0]
        self._tend_back(m, ("tend_h", "tend_u", "tend_h_tl", "tend_u_tl"))
        self._set_diag_state("tlm", h, u, ut)

    def sw_rk4_adj(self):
        """src/mpas_sw_module.f90:783-1002 (one step on the GPU).  On entry
        u(1),h(1) hold the nonlinear state at the step start and
        u_ad(1),h_ad(1) the adjoint at the step end; on return u_ad(1),
        h_ad(1) hold the adjoint at the step start, u(2),h(2) the nonlinear
        end state, and the adjoint levels 2 are zero."""
        m = self._model()
        A = self._arrays
        if "u_ad" not in A:
            raise RuntimeError("var_allocation_adj() has not been called")
        h = self._up(m, A["h"][0, 0], "c")
        u = self._up(m, A["u"][0, 0], "e")
        ha = self._up(m, A["h_ad"][0, 0], "c")
        ua = self._up(m, A["u_ad"][0, 0], "e")
        if any(np.any(A[n + "_ad"] != 0) for n in _RECON):
            # mpas_reconstruct_2d_adj on the end-state adjoint, which consumes
            # the seeds (src/mpas_sw_module.f90:887-892)
            import torch
            seeds = torch.stack([m.to_internal(A[n + "_ad"][0], "c") for n in _RECON])
            ua = ua.to(m.dtype)
            m.reconstruct_adj(seeds, ua)
            ua = ua.to(torch.float64)
            for n in _RECON:
                A[n + "_ad"][...] = 0.0
        seeded = [n for n in _DIAG if n != "vh" and np.any(A[n + "_ad"] != 0)]
        if seeded:
            # sw_compute_solve_diagnostics_adj at the step's end state consumes
            # the seeds of the diagnostic adjoints (src/mpas_sw_module.f90:894-905;
            # vh_ad takes no part, src/mpas_sw_dyncore_adj.f90:294)
            he, ue = h.clone(), u.clone()
            m.base(he, ue, 1)
            kinds = dict(zip(m.lib.msw_diag_names().decode().split(","),
                             m.lib.msw_diag_kinds().decode().split(",")))
            seeds = {}
            for n in seeded:
                g = _GPU_DIAG_NAME.get(n, n)
                seeds[g] = m.to_internal(np.asarray(A[n + "_ad"][0], np.float64), kinds[g])
            m.diagnostics_adj(he, ue, seeds, ha, ua)
            for n in _DIAG:
                if n != "vh":
                    A[n + "_ad"][...] = 0.0
        he, ue = m.adj_step(h, u, ha, ua, want_end=True)
        A["h_ad"][0, 0, :] = m.to_reference(ha, "c")
        A["u_ad"][0, 0, :] = m.to_reference(ua, "e")
        A["h_ad"][1] = 0.0
        A["u_ad"][1] = 0.0
        A["h"][1, 0, :] = m.to_reference(he, "c")
        A["u"][1, 0, :] = m.to_reference(ue, "e")
        for n in ("tend_h_ad", "tend_u_ad"):
            A[n][...] = 0.0
        self._tend_back(m, ("tend_h", "tend_u"))
        self._set_diag_state("adj", h, u)

    def link_tlm_adj(self):
        """Dot-product test of one RK4 step (the reference's
        src/mpas_sw_module.f90:1006-1218 covers RK stage 1 and ends the
        process with STOP; this one covers the whole step, prints the same
        three numbers and returns them)."""
        m = self._model()
        A = self._arrays
        h = self._up(m, A["h"][0, 0], "c")
        u = self._up(m, A["u"][0, 0], "e")
        xh = self._up(m, A["h"][0, 0] * 1e-3, "c")
        xu = self._up(m, A["u"][0, 0] * 1e-3, "e")
        lh, lu = xh.clone(), xu.clone()
        hb, ub = h.clone(), u.clone()
        m.tlm(hb, ub, lh, lu, 1)
        ah, au = lh.clone(), lu.clone()
        m.adj_step(h, u, ah, au)
        lhs = float(m.dot(lh, lu, lh, lu)[0])
        rhs = float(m.dot(xh, xu, ah, au)[0])
        self.lhs, self.rhs = lhs, rhs
        print(lhs, rhs, (lhs - rhs) / lhs)
        return lhs, rhs


# every allocatable array of the module (src/mpas_sw_module.f90:58-201)
_ALL_ARRAYS = set()
for _n in ("u", "h", "tracers"):
    _ALL_ARRAYS |= {_n, _n + "_tl", _n + "_ad"}
_ALL_ARRAYS |= {
    "u_src", "h_s", "fcell", "fvertex", "fedge", "areacell", "invareacell", "dvedge", "dcedge",
    "weightsonedge", "kiteareasonvertex", "meshdensity", "areatriangle", "boundarycell",
    "boundaryedge", "advcells", "cellsoncell", "cellsonedge", "cellsonvertex", "edgesoncell",
    "edgesonedge", "edgesonvertex", "verticesoncell", "verticesonedge", "nedgesoncell",
    "nedgesonedge", "latcell", "loncell", "xcell", "ycell", "zcell", "meshscalingdel2",
    "meshscalingdel4", "deriv_two", "latedge", "lonedge", "xedge", "yedge", "zedge", "angleedge",
    "latvertex", "lonvertex", "xvertex", "yvertex", "zvertex", "defc_a", "defc_b",
    "localverticalunitvectors", "edgenormalvectors", "celltangentplane", "coeffs_reconstruct"}
for _s in ("", "_tl", "_ad"):
    _ALL_ARRAYS |= {n + _s for n in _DIAG + _RECON}
    _ALL_ARRAYS |= {"tend_u" + _s, "tend_h" + _s, "tend_tracer" + _s}

_ROUTINES = ("var_allocation", "var_allocation_tlm", "var_allocation_adj", "sw_mpas_init_block",
             "sw_rk4", "sw_rk4_tlm", "sw_rk4_adj", "link_tlm_adj", "sw_test_case_1",
             "sw_test_case_2", "sw_test_case_5", "sw_test_case_6", "sphere_distance", "aa", "bb", "cc")

mpas_sw_module = _Module()
