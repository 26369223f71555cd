"""Writes tests/golden/task_rollout_layout.json and
tests/golden/policy_pack_layout.json: sizeof and offsetof of every field of
the task-rollout records and of the policy-pack records, as the C compiler
lays them out from a header (the reference's include/metalrobo/c_api.h when
it is present).

The records and their field names are those of the reference header,
c_api.h:55-235 and 334-394; the field lists are read from this repository's
include/mr_c_api.h and include/mr_packs.h, which declare the same records.

Usage: python tests/golden/task_rollout_layout.py [path/to/reference/include]
"""

import json
import os
import re
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
RECORDS = ["MRTaskRolloutDynamicSphereC", "MRTaskRolloutConfigC", "MRTaskRolloutLayoutC", "MRRunManifestC",
           "MRTaskRolloutStageHighWaterC", "MRTaskRolloutAdvanceC"]
FIXTURE = os.path.join(HERE, "task_rollout_layout.json")
PACK_RECORDS = ["MRPolicyDenseLayerC", "MRPolicyPackC", "MRPolicyRolloutBatchC"]
PACK_FIXTURE = os.path.join(HERE, "policy_pack_layout.json")


def record_fields(header, records=RECORDS):
    """Field names of each record, in declaration order."""
    text = re.sub(r"/\*.*?\*/", "", open(header).read(), flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    out = {}
    for name in records:
        m = re.search(r"typedef struct(?: MR_ALIGN16)? " + name + r" \{(.*?)\} " + name + ";", text, re.S)
        if m is None:
            raise RuntimeError(f"{name} not found in {header}")
        fields = []
        for decl in m.group(1).split(";"):
            decl = decl.strip()
            if not decl:
                continue
            fields.append(re.match(r".*?([A-Za-z_][A-Za-z0-9_]*)\s*(\[\d+\])?$", decl, re.S).group(1))
        out[name] = fields
    return out


def layout(include_dir, header_name, fields):
    """sizeof and offsetof of every field, from the C compiler."""
    lines = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{header_name}"', "int main(void) {"]
    for name, fs in fields.items():
        lines.append(f'    printf("{name} size %zu\\n", sizeof({name}));')
        for f in fs:
            lines.append(f'    printf("{name} {f} %zu\\n", offsetof({name}, {f}));')
    lines += ["    return 0;", "}"]
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "layout.c"), os.path.join(d, "layout")
        open(src, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-std=c11", "-I", include_dir, src, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    result = {}
    for line in out.splitlines():
        rec, key, val = line.split()
        result.setdefault(rec, {})[key] = int(val)
    return result


def main():
    ref_include = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/include"
    fields = record_fields(os.path.join(ROOT, "include", "mr_c_api.h"))
    data = layout(ref_include, "metalrobo/c_api.h", fields)
    json.dump({"source": "include/metalrobo/c_api.h:55-235 (LP64, gcc)", "records": data}, open(FIXTURE, "w"),
              indent=1, sort_keys=True)
    print(f"wrote {FIXTURE}")
    fields = record_fields(os.path.join(ROOT, "include", "mr_packs.h"), PACK_RECORDS)
    data = layout(ref_include, "metalrobo/c_api.h", fields)
    json.dump({"source": "include/metalrobo/c_api.h:334-394 (LP64, gcc)", "records": data}, open(PACK_FIXTURE, "w"),
              indent=1, sort_keys=True)
    print(f"wrote {PACK_FIXTURE}")


if __name__ == "__main__":
    main()
