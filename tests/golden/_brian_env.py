"""Set-up for the fixture-writing scripts in this folder (this container only).

Imports brian2 from the read-only reference tree, with its two Cython
extensions compiled from the reference sources by ``make -C oracle ref`` into
``oracle/_ref/brian2ext`` (the reference ships no compiled code and its own
build system is not run).  Nothing in the product or in the GPU tests imports
this module: it only writes the committed fixtures next to it.

Brian is configured for the fp64 verification mode: cython runtime, float64,
and generated code compiled with ``-O2 -ffp-contract=off -fno-fast-math`` in
place of the default ``-O3 -ffast-math -march=native``
(brian2/codegen/cpp_prefs.py:104-135), so that no FMA contraction or
reassociation happens.

It also records, for every hot-path code object, the optimised statement
lists that Brian's code generator produced (brian2/codegen/generators/
cython_generator.py:216-256 receives them in ``translate_statement_sequence``),
together with the variables, indices and conditional writes of the block.
"""
import importlib.abc
import importlib.machinery
import importlib.util
import json
import os
import sys
import sysconfig
import tempfile

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REFBUILD = os.path.join(REPO, "oracle", "_ref")
EXTDIR = os.path.join(REFBUILD, "brian2ext")

_EXT_MODULES = {
    "brian2.synapses.cythonspikequeue": "cythonspikequeue",
    "brian2.memory.cythondynamicarray": "cythondynamicarray",
}


class _RefExtFinder(importlib.abc.MetaPathFinder):
    """Resolves the reference's two compiled extensions to oracle/_ref."""

    def find_spec(self, fullname, path, target=None):
        if fullname not in _EXT_MODULES:
            return None
        fn = os.path.join(EXTDIR, _EXT_MODULES[fullname] + sysconfig.get_config_var("EXT_SUFFIX"))
        if not os.path.exists(fn):
            raise ImportError(f"{fn} missing: run `make -C oracle ref` first")
        loader = importlib.machinery.ExtensionFileLoader(fullname, fn)
        return importlib.util.spec_from_file_location(fullname, fn, loader=loader)


HOT_TEMPLATES = {
    "stateupdate",
    "threshold",
    "reset",
    "synapses",
    "synapses_push_spikes",
    "spikemonitor",
    "statemonitor",
    "spikegenerator",
    "summed_variable",
}

CAPTURED = []


def setup_brian(target="cython"):
    scratch = os.path.join(REFBUILD, "tmp")
    os.makedirs(scratch, exist_ok=True)
    os.environ["TMPDIR"] = scratch
    tempfile.tempdir = scratch
    sys.dont_write_bytecode = True
    if not any(isinstance(f, _RefExtFinder) for f in sys.meta_path):
        sys.meta_path.insert(0, _RefExtFinder())
    if REF not in sys.path:
        sys.path.insert(0, REF)
    import brian2
    from brian2 import prefs

    prefs.codegen.target = target
    prefs.core.default_float_dtype = np.float64
    prefs.codegen.cpp.extra_compile_args_gcc = [
        "-w",
        "-O2",
        "-ffp-contract=off",
        "-fno-fast-math",
        "-std=c++17",
    ]
    prefs.codegen.runtime.cython.cache_dir = os.path.join(REFBUILD, "cython_cache")
    prefs.logging.file_log = False
    prefs.logging.std_redirection = False
    _install_capture()
    return brian2


def _dtype_name(dt):
    if dt is bool or dt is np.bool_:
        return "bool"
    if dt is int:
        return "int64"
    if dt is float:
        return "float64"
    try:
        return np.dtype(dt).name
    except TypeError:
        return str(dt)


def _ser_stmt(s):
    d = {
        "var": s.var,
        "op": s.op,
        "expr": str(s.expr),
        "dtype": _dtype_name(s.dtype),
        "constant": bool(s.constant),
        "subexpression": bool(s.subexpression),
        "scalar": bool(s.scalar),
    }
    if s.used_boolean_variables:
        d["bool_vars"] = list(s.used_boolean_variables)
        d["bool_exprs"] = [
            [[[v, bool(val)] for v, val in key], str(e)]
            for key, e in s.boolean_simplified_expressions.items()
        ]
    return d


def _ser_var(name, var):
    from brian2.core.functions import Function
    from brian2.core.variables import (
        ArrayVariable,
        AuxiliaryVariable,
        Constant,
        DynamicArrayVariable,
        Subexpression,
    )
    from brian2.devices.device import get_device

    e = {"kind": type(var).__name__}
    if isinstance(var, Function):
        e["kind"] = "Function"
        return e
    if hasattr(var, "dtype"):
        e["dtype"] = _dtype_name(var.dtype)
    if isinstance(var, Constant):
        v = var.value
        if isinstance(v, (np.floating, float)):
            v = float(v)
        elif isinstance(v, (np.integer, int, np.bool_, bool)):
            v = int(v) if not isinstance(v, (bool, np.bool_)) else bool(v)
        e["value"] = v
    elif isinstance(var, ArrayVariable):
        e.update(
            size=int(var.size),
            scalar=bool(var.scalar),
            owner=var.owner.name if var.owner is not None else None,
            array_name=get_device().get_array_name(var),
            read_only=bool(var.read_only),
            constant=bool(var.constant),
            dynamic=isinstance(var, DynamicArrayVariable),
        )
    elif isinstance(var, Subexpression):
        e["expr"] = str(var.expr)
    elif isinstance(var, AuxiliaryVariable):
        e["scalar"] = bool(var.scalar)
    return e


def _install_capture():
    from brian2.codegen.generators.cython_generator import CythonCodeGenerator
    from brian2.codegen.generators.numpy_generator import NumpyCodeGenerator

    for cls in (CythonCodeGenerator, NumpyCodeGenerator):
        if getattr(cls, "_hb_captured", False):
            continue
        orig = cls.translate_statement_sequence

        def wrapped(self, sc_statements, ve_statements, _orig=orig):
            if self.template_name in HOT_TEMPLATES:
                blocks = {}
                for bname in sc_statements:
                    sc, ve = sc_statements[bname], ve_statements[bname]
                    sread, swrite, sidx, scw = self.arrays_helper(sc)
                    vread, vwrite, vidx, vcw = self.arrays_helper(ve)
                    blocks[str(bname)] = {
                        "scalar": [_ser_stmt(s) for s in sc],
                        "vector": [_ser_stmt(s) for s in ve],
                        "scalar_read": sorted(sread),
                        "scalar_write": sorted(swrite),
                        "vector_read": sorted(vread),
                        "vector_write": sorted(vwrite),
                        "vector_indices": sorted(vidx),
                        "conditional_write": dict(vcw),
                    }
                used = set()
                for b in blocks.values():
                    for s in b["scalar"] + b["vector"]:
                        used.add(s["var"])
                        from brian2.utils.stringtools import get_identifiers

                        used |= get_identifiers(s["expr"])
                        for _, e in s.get("bool_exprs", []):
                            used |= get_identifiers(e)
                    used |= set(b["scalar_read"]) | set(b["vector_read"]) | set(b["vector_indices"])
                variables = {
                    n: _ser_var(n, v) for n, v in self.variables.items() if n in used
                }
                CAPTURED.append(
                    {
                        "name": self.name,
                        "template": self.template_name,
                        "owner": self.owner.name if self.owner is not None else None,
                        "blocks": blocks,
                        "variables": variables,
                        "indices": {
                            n: self.variable_indices[n]
                            for n in variables
                            if n in self.variable_indices
                        },
                    }
                )
            return _orig(self, sc_statements, ve_statements)

        cls.translate_statement_sequence = wrapped
        cls._hb_captured = True


def take_specs():
    out = list(CAPTURED)
    CAPTURED.clear()
    return out


def dump_json(path, obj):
    with open(path, "w") as f:
        json.dump(obj, f, indent=1, sort_keys=True)
        f.write("\n")
