"""bench.py helpers that run without a GPU: the kernel-source hash that ties
a committed rocprofv3 summary to the kernel it measured."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_code_hash_ignores_comments_and_keeps_literals():
    a = 'int x = 1; // note\n/* block\n comment */ const char *s = "// kept";\nchar q = \'/\';'
    b = 'int x = 1;\nconst char *s = "// kept"; char q = \'/\';   // other note'
    assert bench._code_only(a) == bench._code_only(b)
    assert '"// kept"' in bench._code_only(a) and "'/'" in bench._code_only(a)
    # a change of code changes the text that is hashed
    assert bench._code_only(a) != bench._code_only(a.replace("x = 1", "x = 2"))


def test_kernel_source_hash_is_stable():
    assert bench.kernel_source_sha() == bench.kernel_source_sha()
    assert len(bench.kernel_source_sha()) == 16


def test_kernel_source_hash_covers_device_code_only(tmp_path):
    """A host-launcher edit (below the marker) keeps a profile current; a
    device-code edit makes it stale."""
    dev = "__global__ void k(int *p) { p[0] = 1; }\n"
    host = "// " + bench.HOST_MARKER + "\nint launch() { return 1; }\n"
    a, b, c = tmp_path / "a.hip", tmp_path / "b.hip", tmp_path / "c.hip"
    a.write_text(dev + host)
    b.write_text(dev + host.replace("return 1", "return 2"))
    c.write_text(dev.replace("= 1", "= 2") + host)
    assert bench.kernel_source_sha([str(a)]) == bench.kernel_source_sha([str(b)])
    assert bench.kernel_source_sha([str(a)]) != bench.kernel_source_sha([str(c)])
