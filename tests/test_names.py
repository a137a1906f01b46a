"""Every global name that a function of the repository's Python files reads
is defined at module level or is a builtin (a name error in a rarely run
branch, such as one bench workload, fails here on the CPU instead of on the
GPU box)."""
import builtins
import glob
import os
import symtable

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FILES = sorted(["bench.py", "__graft_entry__.py"] +
               [os.path.relpath(p, ROOT) for d in ("hofstadter_hubbard_tsc_amd", "oracle", "tests")
                for p in glob.glob(os.path.join(ROOT, d, "*.py"))])


def undefined_globals(src, name):
    top = symtable.symtable(src, name, "exec")
    defined = set(top.get_identifiers()) | set(dir(builtins))
    bad = []

    def walk(t):
        for s in t.get_symbols():
            if s.is_global() and s.is_referenced() and s.get_name() not in defined:
                bad.append((t.get_name(), s.get_name()))
        for c in t.get_children():
            walk(c)

    walk(top)
    return bad


@pytest.mark.parametrize("path", FILES)
def test_no_undefined_globals(path):
    with open(os.path.join(ROOT, path)) as f:
        assert undefined_globals(f.read(), path) == []


def test_checker_finds_a_planted_name_error():
    src = "def main():\n    x = 1\n    return x\n\ndef other():\n    return x\n"
    assert undefined_globals(src, "planted.py") == [("other", "x")]
