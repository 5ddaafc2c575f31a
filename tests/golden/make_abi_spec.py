"""Writes ``abi_spec.json``: the reference's native argument lists for the
propagator entry points this library replaces.

Reads the ctypes spec tables of the reference's
``src/tide/backend_utils.py`` as text (the ``_NAME: _Spec = [...]``
assignments, parsed with ``ast`` without importing the module) and expands
them to one ``(kind, name)`` pair per C argument, kind one of ``P``
(pointer), ``F`` (field float), ``I`` (int64), ``B`` (bool), written as
``"K name"`` strings.  The result is
data: the tests compare ``include/tide_hip.h`` and the ctypes bindings with
it.  Run in a tree that holds the reference:

    python tests/golden/make_abi_spec.py [/root/reference]
"""

from __future__ import annotations

import ast
import json
import re
import sys
from pathlib import Path

TEMPLATES = {
    "maxwell_tm_forward": "_TM_FORWARD_SPEC",
    "maxwell_tm_forward_with_storage": "_TM_FORWARD_WITH_STORAGE_SPEC",
    "maxwell_tm_backward": "_TM_BACKWARD_SPEC",
    "maxwell_tm_born_forward": "_TM_BORN_FORWARD_SPEC",
    "maxwell_tm_born_forward_with_storage": "_TM_BORN_FORWARD_WITH_STORAGE_SPEC",
    "maxwell_tm_born_backward_bggrad": "_TM_BORN_BACKWARD_BGGRAD_SPEC",
    "maxwell_3d_forward": "_3D_FORWARD_SPEC",
    "maxwell_3d_forward_with_storage": "_3D_FORWARD_WITH_STORAGE_SPEC",
    "maxwell_3d_backward": "_3D_BACKWARD_SPEC",
    "maxwell_3d_born_forward": "_3D_BORN_FORWARD_SPEC",
}


def main(ref_root: str = "/root/reference") -> None:
    text = (Path(ref_root) / "src/tide/backend_utils.py").read_text()
    tables: dict[str, list] = {}
    for m in re.finditer(r"^(_[A-Z0-9_]+): _Spec = (\[.*?^\])", text, re.S | re.M):
        tables[m.group(1)] = ast.parse(m.group(2), mode="eval").body

    def expand(node) -> list[tuple[str, str]]:
        out: list[tuple[str, str]] = []
        for elt in node.elts:
            if isinstance(elt, ast.Starred):
                out += expand(tables[elt.value.id])
                continue
            kind, count, names = elt.elts
            names = [n.strip() for n in names.value.split(",") if n.strip()]
            if len(names) != count.value:
                raise ValueError(f"count mismatch in {names}")
            out += [(kind.id.lstrip("_"), n) for n in names]
        return out

    spec = {name: [f"{k} {n}" for k, n in expand(tables[var])] for name, var in TEMPLATES.items()}
    out = Path(__file__).resolve().parent / "abi_spec.json"
    out.write_text("{\n" + ",\n".join(f" {json.dumps(k)}: {json.dumps(v)}" for k, v in spec.items())
                   + "\n}\n")
    print(f"wrote {out}: " + ", ".join(f"{k} {len(v)}" for k, v in spec.items()))


if __name__ == "__main__":
    main(*sys.argv[1:])
