"""The identity of the sources the HIP libraries are built from.

`source_hash(root)` is the first 16 hex digits of the SHA-256 over every file that goes into libcospan_hip.so and
libcospan_workloads.so (the C headers, the naga/cospan headers, the workload headers, the .hip sources and their
Makefile), taken in sorted order of their paths relative to the repository root, each as its relative path, a NUL byte,
its bytes and a NUL byte. The Makefile compiles this value into each library (cs_hip_build_id, cs_workloads_build_id),
so a library built from other sources than the tree beside it is told apart before any of its kernels runs.

Run as a script it prints the hash of the repository it sits in (what the Makefile embeds); with file arguments it
hashes those files after the library sources (what a test program built from them embeds).
"""
import glob
import hashlib
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PATTERNS = (
    "include/*.h",
    "cospanmdao_amd/include/naga/cospan/*.hpp",
    "cospanmdao_amd/workloads/*.hpp",
    "cospanmdao_amd/csrc/*.hip",
    "cospanmdao_amd/csrc/*.hpp",
    "cospanmdao_amd/csrc/Makefile",
)


def library_sources(root=ROOT):
    files = set()
    for p in PATTERNS:
        files.update(os.path.relpath(f, root) for f in glob.glob(os.path.join(root, p)))
    return sorted(files)


def source_hash(root=ROOT, extra=()):
    h = hashlib.sha256()
    rels = library_sources(root) + [os.path.relpath(os.path.abspath(e), root) for e in extra]
    for rel in rels:
        h.update(rel.encode() + b"\0")
        with open(os.path.join(root, rel), "rb") as f:
            h.update(f.read())
        h.update(b"\0")
    return h.hexdigest()[:16]


if __name__ == "__main__":
    print(source_hash(ROOT, sys.argv[1:]))
