"""The Boost-free property-tree readers and writers (host/ptree.cpp) under AddressSanitizer and
UndefinedBehaviorSanitizer, on the CPU: every writer's output parses back to the same tree for
generated trees (keys of letters, digits and underscores; values with markup, quotes, escapes,
tabs and newlines inside), and mutated XML / JSON / INFO texts either parse or raise, never
touch memory they do not own.  The readers take user files (hit-event lists, evaluation
settings), so a malformed file must fail cleanly."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HOST = ROOT / "incohautocorrelate_amd" / "host"

PROGRAM = r"""
#include <cctype>
#include <cstdio>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>
#include "ptree.hpp"
using namespace iac_host;

static std::mt19937_64 rng(7);
static bool same(const PTree &a, const PTree &b)
{
    if (a.data != b.data || a.children.size() != b.children.size()) return false;
    for (size_t i = 0; i < a.children.size(); i++)
        if (a.children[i].first != b.children[i].first || !same(a.children[i].second, b.children[i].second))
            return false;
    return true;
}
static std::string key()
{
    const char *a = "abcXYZ_", *b = "abcXYZ_019";
    std::string k(1, a[rng() % 7]);
    for (int i = (int)(rng() % 6); i > 0; i--) k += b[rng() % 10];
    return k;
}
static std::string value()
{
    const char *c = "ab 09<>&\"';{}#\n\t/=.-+e\\";
    std::string d;
    for (int i = (int)(rng() % 8); i > 0; i--) d += c[rng() % 24];
    // the XML reader drops a leaf's surrounding blanks (indented files)
    while (!d.empty() && std::isspace((unsigned char)d.back())) d.pop_back();
    while (!d.empty() && std::isspace((unsigned char)d.front())) d.erase(0, 1);
    return d;
}
static PTree gen(int depth)
{
    PTree t;
    const int n = depth > 0 ? (int)(rng() % 4) : 0;
    if (n == 0) t.data = value();
    for (int i = 0; i < n; i++) t.children.push_back({key(), gen(depth - 1)});
    return t;
}
static PTree parse(int f, const std::string &s)
{
    return f == 0 ? parse_xml(s) : f == 1 ? parse_json(s) : parse_info(s);
}
static std::string write(int f, const PTree &t)
{
    return f == 0 ? write_xml(t) : f == 1 ? write_json(t) : write_info(t);
}

int main()
{
    long bad_round_trips = 0, parsed = 0, rejected = 0;
    for (int it = 0; it < 4000; it++) {
        PTree root;
        root.children.push_back({"root", gen(4)});
        for (int f = 0; f < 3; f++) {
            try {
                if (!same(root, parse(f, write(f, root)))) bad_round_trips++;
            } catch (const std::exception &) {
                bad_round_trips++;
            }
        }
    }
    const std::vector<std::string> seeds = {
        "<?xml version=\"1.0\" encoding=\"utf-8\"?>\n<root><Info><Size>2</Size><Version>9</Version></Info>"
        "<Content><N0><Filename>a.h5</Filename><RotMatrix>1 0 0 0 1 0 0 0 1</RotMatrix></N0>"
        "<N1 a=\"x\">&lt;&amp;&gt;&#65;&#x42;<!-- c --><![CDATA[raw]]></N1></Content></root>\n",
        "{\"root\": {\"Info\": {\"Size\": \"2\"}, \"Content\": {\"N0\": {\"Filename\": \"a\\\"b\\u0041\\n\"},"
        " \"L\": [\"1\", \"2\"], \"x\": 1.5e3, \"t\": true, \"n\": null}}}",
        "root\n{\n  Info\n  {\n    Size 2\n    Version \"9\"\n  }\n  Content\n  {\n    N0 { Filename \"a b\" ; x\n"
        " Event 3 }\n  }\n}\n",
    };
    const char alphabet[] = "<>/=\"'&;#{}[]:,\\ \n\tabcxyz019!?-.uU";
    for (long it = 0; it < 30000; it++) {
        const int f = (int)(rng() % 3);
        std::string s = seeds[f];
        for (int m = 1 + (int)(rng() % 8); m > 0 && !s.empty(); m--) {
            const size_t pos = rng() % (s.size() + 1);
            switch (rng() % 4) {
            case 0: if (pos < s.size()) s.erase(pos, 1 + rng() % 4); break;
            case 1: s.insert(pos, 1, alphabet[rng() % (sizeof(alphabet) - 1)]); break;
            case 2: if (pos < s.size()) s[pos] = alphabet[rng() % (sizeof(alphabet) - 1)]; break;
            default: s.resize(pos); break;
            }
        }
        try {
            const PTree t = parse(f, s);
            for (int w = 0; w < 3; w++) {
                try {
                    (void)parse(w, write(w, t));
                } catch (const std::exception &) {
                }
            }
            parsed++;
        } catch (const std::exception &) {
            rejected++;
        }
    }
    std::printf("%ld %ld %ld\n", bad_round_trips, parsed, rejected);
    return 0;
}
"""


def test_ptree_round_trips_and_mutations_under_sanitizers(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    src = tmp_path / "fuzz.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "fuzz"
    build = subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
                            "-fno-sanitize-recover=undefined", "-I", str(HOST), str(src),
                            str(HOST / "ptree.cpp"), "-o", str(exe)],
                           capture_output=True, text=True, timeout=300)
    if build.returncode != 0 and "sanitizer" in build.stderr.lower():
        pytest.skip("the host compiler has no sanitizer runtime")
    assert build.returncode == 0, build.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    bad, parsed, rejected = (int(x) for x in r.stdout.split())
    assert bad == 0
    assert parsed > 1000 and rejected > 1000  # the mutations reach both outcomes
