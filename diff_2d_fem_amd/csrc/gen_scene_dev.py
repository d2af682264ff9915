"""Write scene_dev.h: a device-side mirror of d2f_scene (include/d2fem.h) whose pointer fields
carry the global address space.

The scene kernel reads its d2f_scene from an LDS copy. A pointer loaded from memory has no
address space the compiler can prove, so every access through it would be a generic (flat)
instruction; flat instructions count against both the vector-memory and the LDS wait
counters, which serializes the kernel's LDS work behind its global loads. Reading the same
bytes through this mirror makes every field access a global one. Field order, types and
offsets are the header's (checked by the static_asserts written alongside).

    python diff_2d_fem_amd/csrc/gen_scene_dev.py      (build.py runs it before compiling)
"""
from __future__ import annotations

import re
from pathlib import Path

HERE = Path(__file__).resolve().parent
HEADER = HERE.parent.parent / "include" / "d2fem.h"
OUT = HERE / "scene_dev.h"


def fields(src):
    body = re.search(r"typedef struct d2f_scene \{(.*?)\} d2f_scene;", src, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", " ", body, flags=re.S)
    out = []
    for stmt in body.split(";"):
        stmt = " ".join(stmt.split())
        if not stmt:
            continue
        m = re.match(r"^((?:const )?[A-Za-z_][A-Za-z_0-9]*)\s+(.*)$", stmt)
        base, decls = m.group(1), m.group(2)
        for d in decls.split(","):
            d = d.strip()
            ptr = d.startswith("*")
            name = d.lstrip("*").strip()
            out.append((base, ptr, name))
    return out


def render(flds):
    lines = [
        "// Generated by gen_scene_dev.py from include/d2fem.h: do not edit.",
        "// d2f_scene with every pointer field in the global address space (see the generator).",
        "#pragma once",
        "#include <cstddef>",
        "",
        '#include "../../include/d2fem.h"',
        "",
        "#ifndef D2F_GLOBAL",
        "#define D2F_GLOBAL __attribute__((address_space(1)))",
        "#endif",
        "",
        "namespace d2f {",
        "",
        "struct SceneG {",
    ]
    for base, ptr, name in flds:
        if ptr:
            cv, t = ("const ", base[6:]) if base.startswith("const ") else ("", base)
            lines.append(f"  {cv}D2F_GLOBAL {t} *{name};")
        else:
            lines.append(f"  {base} {name};")
    lines += ["};", "", "static_assert(sizeof(SceneG) == sizeof(d2f_scene), \"SceneG size\");"]
    for _, _, name in flds:
        fld = name.split("[")[0]
        lines.append(f"static_assert(offsetof(SceneG, {fld}) == offsetof(d2f_scene, {fld}), \"{fld}\");")
    lines += ["", "}  // namespace d2f", ""]
    return "\n".join(lines)


def generate():
    text = render(fields(HEADER.read_text()))
    if not OUT.exists() or OUT.read_text() != text:
        OUT.write_text(text)
    return OUT


if __name__ == "__main__":
    print(generate())
