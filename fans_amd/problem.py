"""Problem description: the reference's JSON input format, parsed host-side.

Restates src/reader.cpp:63-183 (Reader::ReadInputFile), include/mixedBCs.h:51-133
(MixedBC::from_json) and include/MaterialManager.h:38-110 (material groups,
legacy top-level "matmodel").  The microstructure is an array of phase ids in
x-major (x, y, z) order, like Reader::ms after ReadMS (reader.cpp:380-399).
"""
from __future__ import annotations

import json
import os
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

# Model names accepted by createMatmodel (include/setup.h:21-61) that this
# implementation runs, with the material_properties keys in parameter order.
MODEL_PARAMS = {
    "LinearThermalIsotropic": ("conductivity",),
    "LinearThermalTriclinic": ("K_11", "K_12", "K_13", "K_22", "K_23", "K_33"),
    "LinearElasticIsotropic": ("bulk_modulus", "shear_modulus"),
    "LinearElasticTriclinic": tuple(f"C_{i}{j}" for i in range(1, 7) for j in range(i, 7)),
    "J2PlasticityNew_LinearIsotropicHardening": ("bulk_modulus", "shear_modulus", "yield_stress",
                                                 "isotropic_hardening_parameter"),
    "PseudoPlasticLinearHardening": ("bulk_modulus", "shear_modulus", "yield_stress", "hardening_parameter"),
    "PseudoPlasticNonLinearHardening": ("bulk_modulus", "shear_modulus", "yield_stress", "hardening_exponent",
                                        "eps_0"),
    "SaintVenantKirchhoff": ("bulk_modulus", "shear_modulus"),
    "CompressibleNeoHookean": ("bulk_modulus", "shear_modulus"),
    "J2ViscoPlastic_LinearIsotropicHardening": ("bulk_modulus", "shear_modulus", "yield_stress",
                                                "isotropic_hardening_parameter", "kinematic_hardening_parameter",
                                                "viscosity", "time_step"),
    "J2ViscoPlastic_NonLinearIsotropicHardening": ("bulk_modulus", "shear_modulus", "yield_stress",
                                                   "isotropic_hardening_parameter", "kinematic_hardening_parameter",
                                                   "viscosity", "time_step", "saturation_stress",
                                                   "saturation_exponent"),
}
# keys the reference reads as one number for the whole model (J2Plasticity.h:19);
# every phase of the group carries a copy
SCALAR_PARAMS = {"time_step"}
MODEL_KIND = {
    "LinearThermalIsotropic": 1,
    "LinearThermalTriclinic": 2,
    "LinearElasticIsotropic": 3,
    "LinearElasticTriclinic": 4,
    "J2PlasticityNew_LinearIsotropicHardening": 5,
    "PseudoPlasticLinearHardening": 6,
    "SaintVenantKirchhoff": 7,
    "CompressibleNeoHookean": 8,
    "PseudoPlasticNonLinearHardening": 9,
    "J2ViscoPlastic_LinearIsotropicHardening": 10,
    "J2ViscoPlastic_NonLinearIsotropicHardening": 11,
}
# createMatmodel<3,9> (include/setup.h:62-73): finite strain, F and P row major
LARGE_STRAIN_MODELS = {"SaintVenantKirchhoff", "CompressibleNeoHookean"}
THERMAL_MODELS = {"LinearThermalIsotropic", "LinearThermalTriclinic"}
LINEAR_MODELS = {"LinearThermalIsotropic", "LinearThermalTriclinic", "LinearElasticIsotropic",
                 "LinearElasticTriclinic"}


@dataclass
class MaterialGroup:
    matmodel: str
    phases: list
    properties: dict

    def param_matrix(self) -> np.ndarray:
        """[n_phases, n_params] in MODEL_PARAMS order."""
        if self.matmodel not in MODEL_PARAMS:
            raise ValueError(f"{self.matmodel} is not a supported matmodel")
        keys = MODEL_PARAMS[self.matmodel]
        cols = []
        for k in keys:
            if k not in self.properties:
                raise ValueError(f"Missing material property '{k}' for {self.matmodel}")
            v = np.atleast_1d(np.asarray(self.properties[k], dtype=np.float64))
            if k in SCALAR_PARAMS:
                if np.ndim(self.properties[k]) != 0:
                    raise ValueError(f"Material property '{k}' must be a number")
                v = np.full(len(self.phases), float(v[0]))
            if v.size != len(self.phases):
                raise ValueError(f"Inconsistent size for material property: {k}")
            cols.append(v)
        return np.stack(cols, axis=1)


@dataclass
class LoadCase:
    """include/mixedBCs.h:139-144."""
    mixed: bool
    g0_path: Optional[np.ndarray] = None      # [n_steps, n_str] (pure strain)
    mbc: Optional[dict] = None                # strain_indices, stress_indices, strain_path, stress_path, n_steps
    n_steps: int = 0


@dataclass
class Problem:
    dims: tuple
    howmany: int
    ms: np.ndarray
    groups: list
    L: tuple = (1.0, 1.0, 1.0)
    fe_type: str = "HEX8"
    method: str = "cg"
    measure: str = "Linfinity"
    err_type: str = "absolute"
    tol: float = 1e-10
    n_it: int = 100
    ls_max_iter: int = 5
    ls_tol: float = 1e-2
    ref_material: Optional[np.ndarray] = None
    load_cases: list = field(default_factory=list)
    results: list = field(default_factory=list)
    results_prefix: str = ""
    extrapolate_displacement: bool = True
    strain_type: str = "small"
    problem_type: str = "mechanical"
    ms_filename: str = ""
    ms_datasetname: str = ""

    @property
    def n_str(self) -> int:
        if self.howmany == 1:
            return 3
        return 9 if self.strain_type == "large" else 6

    @property
    def large_strain(self) -> bool:
        return self.howmany == 3 and self.strain_type == "large"

    @property
    def all_linear(self) -> bool:
        return all(g.matmodel in LINEAR_MODELS for g in self.groups)

    def validate(self):
        nx, ny, nz = self.dims
        if self.ms.shape != (nx, ny, nz):
            raise ValueError("microstructure shape does not match dims")
        if nz % 2 != 0:
            raise ValueError("n_z must be even")
        for g in self.groups:
            thermal = g.matmodel in THERMAL_MODELS
            if thermal != (self.howmany == 1):
                raise ValueError(f"{g.matmodel} is not a valid matmodel for this problem type")
            if self.howmany == 3 and (g.matmodel in LARGE_STRAIN_MODELS) != self.large_strain:
                kind = "large" if self.large_strain else "small"
                raise ValueError(f"{g.matmodel} is not a valid {kind} strain material model")
            g.param_matrix()  # missing / inconsistent material properties raise here
        seen = set()
        for g in self.groups:
            for p in g.phases:
                if p < 0 or p in seen:
                    raise ValueError(f"MaterialManager: Invalid or duplicate phase {p}")
                seen.add(p)
        n_phases = max(seen) + 1
        for p in range(n_phases):
            if p not in seen:
                raise ValueError(f"MaterialManager: Phase {p} not assigned")
        if int(self.ms.max()) >= n_phases:
            raise ValueError("microstructure holds a phase id without material")

    def flat_materials(self):
        """(kinds, nphase, phase_ids, params) as int32/float64 arrays."""
        kinds = np.array([MODEL_KIND[g.matmodel] for g in self.groups], dtype=np.int32)
        nph = np.array([len(g.phases) for g in self.groups], dtype=np.int32)
        pids = np.array([p for g in self.groups for p in g.phases], dtype=np.int32)
        params = np.concatenate([g.param_matrix().reshape(-1) for g in self.groups]).astype(np.float64)
        return kinds, nph, pids, params

    # ------------------------------------------------------------------ JSON
    @staticmethod
    def from_json(src, ms: Optional[np.ndarray] = None, base_dir: Optional[str] = None) -> "Problem":
        """Parse the reference input format (reader.cpp:63-183, README.md:240-364)."""
        if isinstance(src, (str, os.PathLike)):
            base_dir = base_dir or os.path.dirname(os.path.abspath(src))
            with open(src) as f:
                j = json.load(f)
        else:
            j = src
        mi = j["microstructure"]
        dsname = mi["datasetname"]
        if not dsname:
            raise ValueError("datasetname must not be empty and must refer to a valid HDF5 path")
        if dsname[0] != "/":
            dsname = "/" + dsname
        problem_type = j["problem_type"]
        strain_type = j.get("strain_type", "small")
        if strain_type not in ("small", "large"):
            raise ValueError("strain_type must be either 'small' or 'large'")
        if problem_type == "thermal":
            howmany, n_str = 1, 3
        elif problem_type == "mechanical" and strain_type == "small":
            howmany, n_str = 3, 6
        elif problem_type == "mechanical" and strain_type == "large":
            howmany, n_str = 3, 9
        else:
            raise ValueError(f"{problem_type}/{strain_type} is not a supported problem type")
        fe_type = j.get("FE_type", "HEX8")
        if fe_type not in ("HEX8", "HEX8R", "BBAR"):
            raise ValueError("FE_type must be one of: 'HEX8', 'HEX8R', or 'BBAR'")
        ep = j["error_parameters"]
        ls = j.get("linesearch_parameters", {})
        ls_max_iter = int(ls.get("max_iter", 5))
        ls_tol = float(ls.get("tol", 1e-2))
        if ls_max_iter < 1 or ls_tol <= 0.0:
            raise ValueError("linesearch_parameters: max_iter >= 1 and tol > 0 required")
        # material groups (MaterialManager.h:40-65, legacy format included)
        if "matmodel" in j and "material_properties" in j:
            props = j["material_properties"]
            n_mats = next(len(v) for v in props.values() if isinstance(v, list))
            mats = [dict(matmodel=j["matmodel"], material_properties=props, phases=list(range(n_mats)))]
        else:
            mats = j["materials"]
        if not isinstance(mats, list) or not mats:
            raise ValueError("MaterialManager: 'materials' must be non-empty array")
        groups = []
        for mg in mats:
            if not all(k in mg for k in ("phases", "matmodel", "material_properties")):
                raise ValueError("MaterialManager: material group missing required fields")
            groups.append(MaterialGroup(mg["matmodel"], [int(p) for p in mg["phases"]], mg["material_properties"]))
        ref = None
        if "reference_material" in j:
            ref = np.asarray(j["reference_material"], dtype=np.float64)
            if ref.shape != (n_str, n_str):
                raise ValueError(f"reference_material must be {n_str}x{n_str}")
            try:
                np.linalg.cholesky(ref)
            except np.linalg.LinAlgError:
                raise ValueError("reference_material must be symmetric positive definite")
        load_cases = []
        ml = j["macroscale_loading"]
        if not isinstance(ml, list):
            raise ValueError("macroscale_loading must be an array")
        for entry in ml:
            if isinstance(entry, list):
                path = np.asarray(entry, dtype=np.float64)
                if path.ndim != 2 or path.shape[1] != n_str:
                    raise ValueError(f"Invalid length of loading vector: expected {n_str} components")
                load_cases.append(LoadCase(mixed=False, g0_path=path, n_steps=path.shape[0]))
            else:
                mbc = mixed_bc_from_json(entry, n_str)
                load_cases.append(LoadCase(mixed=True, mbc=mbc, n_steps=mbc["n_steps"]))
        if ms is None:
            from . import hdf5io
            fp = mi["filepath"]
            if base_dir and not os.path.isabs(fp):
                fp = os.path.join(base_dir, fp)
            ms = hdf5io.read_microstructure(fp, dsname)
        ms = np.ascontiguousarray(ms, dtype=np.uint16)
        p = Problem(dims=tuple(int(d) for d in ms.shape), howmany=howmany, ms=ms, groups=groups,
                    L=tuple(float(v) for v in mi["L"]), fe_type=fe_type, method=j["method"],
                    measure=ep["measure"], err_type=ep["type"], tol=float(ep["tolerance"]), n_it=int(j["n_it"]),
                    ls_max_iter=ls_max_iter, ls_tol=ls_tol, ref_material=ref, load_cases=load_cases,
                    results=list(j.get("results", [])), results_prefix=j.get("results_prefix", ""),
                    extrapolate_displacement=bool(j.get("extrapolate_displacement", True)),
                    strain_type=strain_type, problem_type=problem_type, ms_filename=mi["filepath"],
                    ms_datasetname=dsname)
        if p.method not in ("cg", "fp"):
            raise ValueError(f"{p.method} is not a valid method")
        if p.measure not in ("L1", "L2", "Linfinity"):
            raise ValueError(f"Unknown measure type: {p.measure}")
        if p.err_type not in ("absolute", "relative"):
            raise ValueError(f"Unknown error type: {p.err_type}")
        p.validate()
        return p


def mixed_bc_from_json(jc: dict, n_str: int) -> dict:
    """MixedBC::from_json (include/mixedBCs.h:51-133)."""
    if "strain_indices" not in jc or "stress_indices" not in jc:
        raise ValueError("mixed BC: strain_indices or stress_indices missing")
    iE = [int(v) for v in jc["strain_indices"]]
    iF = [int(v) for v in jc["stress_indices"]]
    present = [0] * n_str
    for k in iE:
        if k < 0 or k >= n_str:
            raise ValueError("strain index out of range")
        present[k] = 1
    for k in iF:
        if k < 0 or k >= n_str:
            raise ValueError("stress index out of range")
        if present[k]:
            raise ValueError("index appears in both strain_indices and stress_indices")
        present[k] = 1
    if not all(present):
        raise ValueError("each component must be either strain- or stress-controlled")
    strain_raw, stress_raw, n_steps = [], [], 0
    if iE:
        if "strain" not in jc:
            raise ValueError("strain array missing")
        strain_raw = jc["strain"]
        n_steps = len(strain_raw)
    if iF:
        if "stress" not in jc:
            raise ValueError("stress array missing")
        stress_raw = jc["stress"]
        n_steps = max(n_steps, len(stress_raw))
    if n_steps == 0:
        raise ValueError("mixed BC: at least one of strain/stress must have timesteps")
    if not strain_raw:
        strain_raw = [[] for _ in range(n_steps)]
    if not stress_raw:
        stress_raw = [[] for _ in range(n_steps)]
    FE = np.zeros((n_steps, len(iE)))
    PF = np.zeros((n_steps, len(iF)))
    for t in range(n_steps):
        if len(strain_raw[t]) != len(iE):
            raise ValueError("strain row length mismatch")
        if len(stress_raw[t]) != len(iF):
            raise ValueError("stress row length mismatch")
        FE[t, :] = strain_raw[t]
        PF[t, :] = stress_raw[t]
    return dict(strain_indices=iE, stress_indices=iF, strain_path=FE, stress_path=PF, n_steps=n_steps)


def isotropic_KG(E: float, nu: float):
    """Bulk and shear modulus from Young's modulus and Poisson's ratio."""
    return E / (3.0 * (1.0 - 2.0 * nu)), E / (2.0 * (1.0 + nu))
