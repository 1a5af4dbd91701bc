"""The Boost-free configuration files of the host mirror (CPU only): the hit-event list
(Settings::LoadHitEventListFromFile / SafeHitEventListToFile, Settings.cpp:473-676) and the
evaluation settings (MainRunModes::LoadEvaluationSettings, MainRunModes.cpp:32-91).

A small C++ program against libiac_host.so loads a file and prints what it read as JSON.
The fixtures are written in the reference's layouts: root.Info.Size / Version and
root.Content.N<i> as boost's write_xml (one line), an indented XML, write_json (every value a
string) and write_info produce them; the version-1 R0..R8 keys; the oldest root.content.<i>
lists; rotation strings with too many entries; entities; missing optional keys."""
import json
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "incohautocorrelate_amd"

PROGRAM = r"""
#include <cstdio>
#include <cstring>
#include <string>
#include "iac_host.hpp"

static std::string esc(const std::string &s)
{
    std::string o;
    for (char c : s) {
        if (c == '"' || c == '\\') o += '\\';
        o += c;
    }
    return o;
}

int main(int argc, char **argv)
{
    const std::string mode = argv[1];
    Settings S;
    S.echo = false;
    if (mode == "hitlist" || mode == "roundtrip") {
        if (mode == "roundtrip") {
            for (int i = 0; i < 3; i++) {
                Settings::HitEvent e;
                e.Filename = i == 1 ? "a&b <c> \"d\".h5" : "stack_" + std::to_string(i) + ".bin";
                e.Dataset = "/data/frames";
                e.Event = 7 * i;
                e.SerialNumber = 100 + i;
                e.MeanIntensity = 0.1f * (float)(i + 1);
                e.PhotonCount = 2000 + i;
                e.SupplementInfo = i == 2 ? "" : "run 17";
                for (int j = 0; j < 9; j++) e.RotMatrix[j] = (float)(j + 1) / 7.0f - (float)i;
                S.HitEvents.push_back(e);
            }
            S.SafeHitEventListToFile(argv[2]);
            S.HitEvents.clear();
        }
        S.LoadHitEventListFromFile(argv[2]);
        std::printf("[");
        for (size_t i = 0; i < S.HitEvents.size(); i++) {
            const auto &e = S.HitEvents[i];
            std::printf("%s{\"Filename\": \"%s\", \"Dataset\": \"%s\", \"Event\": %d, \"SerialNumber\": %d, "
                        "\"MeanIntensity\": %.9g, \"PhotonCount\": %d, \"SupplementInfo\": \"%s\", \"Rot\": [",
                        i ? ", " : "", esc(e.Filename).c_str(), esc(e.Dataset).c_str(), e.Event, e.SerialNumber,
                        (double)e.MeanIntensity, e.PhotonCount, esc(e.SupplementInfo).c_str());
            for (int j = 0; j < 9; j++) std::printf("%s%.9g", j ? ", " : "", (double)e.RotMatrix[j]);
            std::printf("]}");
        }
        std::printf("]\n");
        return 0;
    }
    if (mode == "eval") {
        RunIAC::CreateDataEval_Settings E = MainRunModes::LoadEvaluationSettings(argv[2], S);
        std::printf("{\"XML_Path\": \"%s\", \"PixelMap_Path\": \"%s\", \"PixelMap_DataSet\": \"%s\", "
                    "\"PixelMask_Path\": \"%s\", \"PixelMask_Dataset\": \"%s\", \"UseExistingAvInt\": %d, "
                    "\"Out_Final_AC_Path\": \"%s\", \"Out_AvIntensity_Path\": \"%s\", \"Out_Cq_Path\": \"%s\", "
                    "\"Out_Cq_small_Path\": \"%s\", \"Out_ACuw_Path\": \"%s\", \"Out_Q_Vector\": \"%s\", "
                    "\"MeshSize\": %d, \"QZoom\": %.9g, \"AngularAveraged\": %d, \"DoubleMap\": %d, "
                    "\"Interp1\": %d, \"Interp2\": %d, \"PhotonOffset\": %.9g, \"PhotonStep\": %.9g, "
                    "\"RestrictStackToBoundaries\": %d, \"LowerBoundary\": %u, \"UpperBoundary\": %u, "
                    "\"UsePixelMask_as_Flatfield\": %d, \"FractionalCq\": %d, \"SizeOfCqFraction\": %u, "
                    "\"InvertRotMatrix\": %d, \"EchoLevel\": %d, \"Out_Report\": \"%s\", \"DetectorSize\": [%u, %u], "
                    "\"echo\": %d}\n",
                    E.XML_Path.c_str(), E.PixelMap_Path.c_str(), E.PixelMap_DataSet.c_str(), E.PixelMask_Path.c_str(),
                    E.PixelMask_Dataset.c_str(), (int)E.UseExistingAvInt, E.Out_Final_AC_Path.c_str(),
                    E.Out_AvIntensity_Path.c_str(), E.Out_Cq_Path.c_str(), E.Out_Cq_small_Path.c_str(),
                    E.Out_ACuw_Path.c_str(), E.Out_Q_Vector.c_str(), E.MeshSize, (double)E.QZoom,
                    (int)E.AngularAveraged, (int)E.DoubleMap, (int)E.AC_FirstMap_Flags.InterpolationMode,
                    (int)E.AC_SecondMap_Flags.InterpolationMode, (double)E.PhotonOffset, (double)E.PhotonStep,
                    (int)E.RestrictStackToBoundaries, E.LowerBoundary, E.UpperBoundary,
                    (int)E.UsePixelMask_as_Flatfield, (int)E.FractionalCq, E.SizeOfCqFraction,
                    (int)E.InvertRotMatrix, E.EchoLevel, E.Out_Report.c_str(), E.DetectorSize[0], E.DetectorSize[1],
                    (int)S.echo);
        return 0;
    }
    if (mode == "example") return MainRunModes::Create_Example_Evaluation_Config_File(argv[2], S);
    return 2;
}
"""


@pytest.fixture(scope="module")
def reader(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    if not (PKG / "libiac_host.so").exists():
        pytest.skip("host library not built")
    d = tmp_path_factory.mktemp("cfg")
    src = d / "reader.cpp"
    src.write_text(PROGRAM)
    exe = d / "reader"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", str(ROOT / "include"), str(src), "-o", str(exe),
                    "-L", str(PKG), "-liac_host", "-liac_hip", f"-Wl,-rpath,{PKG}", "-Wl,-rpath,/opt/rocm/lib"],
                   check=True, capture_output=True, text=True)

    def run(mode, path):
        r = subprocess.run([str(exe), mode, str(path)], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stdout + r.stderr
        return json.loads(r.stdout), r.stderr

    return run


ROT = "0.000000; -1.000000; 0.000000; 1.000000; 0.000000; 0.000000; 0.000000; 0.000000; 1.000000"
EVENTS = [
    dict(Filename="/data/run17/stack.bin", Dataset="/entry/data", Event=3, SerialNumber=11, MeanIntensity="0.100000001",
         PhotonCount=2095, SupplementInfo="gain corrected", Rotation=ROT),
    dict(Filename="a&amp;b.bin", Dataset="", Event=0, SerialNumber=12, MeanIntensity="-1", PhotonCount=-1,
         Rotation="1; 0; 0; 0; 1; 0; 0; 0; 1"),
]


def _xml_one_line(events, version=9, content="Content", prefix="N"):
    body = "".join(f"<{prefix}{i}>" + "".join(f"<{k}>{v}</{k}>" for k, v in e.items()) + f"</{prefix}{i}>"
                   for i, e in enumerate(events))
    return (f'<?xml version="1.0" encoding="utf-8"?>\n<root><Info><Size>{len(events)}</Size>'
            f"<Version>{version}</Version></Info><{content}>{body}</{content}></root>")


def _check_events(got):
    assert len(got) == 2
    a, b = got
    assert a["Filename"] == "/data/run17/stack.bin" and a["Dataset"] == "/entry/data"
    assert (a["Event"], a["SerialNumber"], a["PhotonCount"]) == (3, 11, 2095)
    assert np.float32(a["MeanIntensity"]) == np.float32(0.1)
    assert a["SupplementInfo"] == "gain corrected"
    assert a["Rot"] == [0, -1, 0, 1, 0, 0, 0, 0, 1]
    assert b["Filename"] == "a&b.bin" and b["Dataset"] == ""
    assert b["MeanIntensity"] == -1 and b["PhotonCount"] == -1
    assert b["SupplementInfo"] == ""  # optional key, absent
    assert b["Rot"] == [1, 0, 0, 0, 1, 0, 0, 0, 1]


def test_hitlist_boost_xml(reader, tmp_path):
    f = tmp_path / "events.xml"
    f.write_text(_xml_one_line(EVENTS))
    _check_events(reader("hitlist", f)[0])


def test_hitlist_indented_xml_with_comment(reader, tmp_path):
    xml = _xml_one_line(EVENTS).replace("><", ">\n    <").replace("<root>", "<!-- made by hand -->\n<root>")
    f = tmp_path / "events.XML"
    f.write_text(xml)
    _check_events(reader("hitlist", f)[0])


def test_hitlist_json_and_info(reader, tmp_path):
    def jnode(e):
        return {k: str(v).replace("&amp;", "&") for k, v in e.items()}
    doc = {"root": {"Info": {"Size": "2", "Version": "9"},
                    "Content": {f"N{i}": jnode(e) for i, e in enumerate(EVENTS)}}}
    f = tmp_path / "events.json"
    f.write_text(json.dumps(doc, indent=4))
    _check_events(reader("hitlist", f)[0])
    lines = ["root", "{", "    Info", "    {", "        Size 2", "        Version 9", "    }", "    Content", "    {"]
    for i, e in enumerate(EVENTS):
        lines += [f"        N{i}", "        {"]
        for k, v in e.items():
            v = str(v).replace("&amp;", "&")
            lines.append(f'            {k} "{v}"' if (" " in v or v == "") else f"            {k} {v}")
        lines += ["        }"]
    lines += ["    }", "}", "; a comment"]
    f = tmp_path / "events.info"
    f.write_text("\n".join(lines) + "\n")
    _check_events(reader("hitlist", f)[0])


def test_hitlist_version1_and_oldest_layouts(reader, tmp_path):
    ev = dict(EVENTS[0])
    del ev["Rotation"]
    ev.update({f"R{j}": v for j, v in enumerate([0, 0, 1, 0, 1, 0, -1, 0, 0])})
    f = tmp_path / "v1.xml"
    f.write_text(_xml_one_line([ev], version=1, content="content"))
    got = reader("hitlist", f)[0]
    assert got[0]["Rot"] == [0, 0, 1, 0, 1, 0, -1, 0, 0]
    ev2 = dict(EVENTS[0])
    del ev2["Rotation"]
    f = tmp_path / "old.xml"  # root.content.<i>, no Version: version 1 keys, identity by default
    f.write_text(_xml_one_line([ev2], version=1, content="content", prefix="").replace("<Version>1</Version>", ""))
    got = reader("hitlist", f)[0]
    assert got[0]["Rot"] == [1, 0, 0, 0, 1, 0, 0, 0, 1] and got[0]["Event"] == 3


def test_hitlist_bad_rotation_is_identity(reader, tmp_path):
    ev = dict(EVENTS[0], Rotation="1; 2; 3; 4; 5; 6; 7; 8; 9; 10")
    f = tmp_path / "bad.xml"
    f.write_text(_xml_one_line([ev]))
    got, err = reader("hitlist", f)
    assert got[0]["Rot"] == [1, 0, 0, 0, 1, 0, 0, 0, 1]
    assert "Invalid rotation matrix format" in err


@pytest.mark.parametrize("ext", ["xml", "json", "info"])
def test_hitlist_round_trip(reader, tmp_path, ext):
    got = reader("roundtrip", tmp_path / f"rt.{ext}")[0]
    assert [e["Event"] for e in got] == [0, 7, 14]
    assert got[1]["Filename"] == 'a&b <c> "d".h5'
    assert got[0]["SupplementInfo"] == "run 17" and got[2]["SupplementInfo"] == ""
    for i, e in enumerate(got):
        assert np.float32(e["MeanIntensity"]) == np.float32(0.1) * np.float32(i + 1)
        # the rotation goes through std::to_string (six decimals), as the reference writes it
        want = [float(f"{float(np.float32(np.float32(j + 1) / np.float32(7.0) - np.float32(i))):.6f}")
                for j in range(9)]
        assert np.array_equal(np.float32(e["Rot"]), np.float32(want))
    xml = (tmp_path / "rt.xml").read_text() if ext == "xml" else None
    if xml:
        assert "<Version>9</Version>" in xml and "<N2>" in xml and "&amp;" in xml


def test_evaluation_settings(reader, tmp_path):
    full = """<?xml version="1.0" encoding="utf-8"?>
<root><Info><Version>12</Version></Info><EvalSettings>
  <InputFiles><EventList_Path>ev.xml</EventList_Path><PixelMap_Path>pm.bin</PixelMap_Path>
    <PixelMap_Dataset>/pm</PixelMap_Dataset><PixelMask_Path>mask.bin</PixelMask_Path>
    <PixelMask_Dataset>/mask</PixelMask_Dataset><UseExistingAvInt>true</UseExistingAvInt>
    <PixelMap_Size>512; 256</PixelMap_Size></InputFiles>
  <OutputFiles><AC_Final_Path>ac.bin</AC_Final_Path><AvIntensity_Path>av.bin</AvIntensity_Path>
    <Cq_Merged_Path>cq.bin</Cq_Merged_Path><Cq_small_Path>cqs.bin</Cq_small_Path><ACuw_Path>acuw.bin</ACuw_Path>
    <Q_Vector>q.bin</Q_Vector></OutputFiles>
  <OutputMesh><MeshSize>255</MeshSize><QZoom>1.5</QZoom><AngularAveraged>1</AngularAveraged>
    <DoubleMap>false</DoubleMap><AC_FirstMap_Flags><InterpolationMode>1</InterpolationMode></AC_FirstMap_Flags>
    <AC_SecondMap_Flags><InterpolationMode>1</InterpolationMode></AC_SecondMap_Flags></OutputMesh>
  <PhotonThreshold><PhotonOffset>0.5</PhotonOffset><PhotonStep>2.5</PhotonStep></PhotonThreshold>
  <Misc><RestrictStackToBoundaries>true</RestrictStackToBoundaries><LowerBoundary>5</LowerBoundary>
    <UpperBoundary>500</UpperBoundary><UsePixelmaskAsFlatField>true</UsePixelmaskAsFlatField>
    <FractionalCq>false</FractionalCq><SizeOfCqFraction>20</SizeOfCqFraction>
    <InvertRotMatrix>true</InvertRotMatrix><EchoLevel>0</EchoLevel><Report_Path>rep.txt</Report_Path></Misc>
</EvalSettings></root>
"""
    f = tmp_path / "full.xml"
    f.write_text(full)
    e = reader("eval", f)[0]
    assert (e["XML_Path"], e["PixelMap_Path"], e["PixelMap_DataSet"], e["PixelMask_Path"]) == \
        ("ev.xml", "pm.bin", "/pm", "mask.bin")
    assert e["UseExistingAvInt"] == 1 and e["DetectorSize"] == [512, 256]
    assert (e["Out_Final_AC_Path"], e["Out_AvIntensity_Path"], e["Out_Cq_Path"], e["Out_Cq_small_Path"],
            e["Out_ACuw_Path"], e["Out_Q_Vector"]) == ("ac.bin", "av.bin", "cq.bin", "cqs.bin", "acuw.bin", "q.bin")
    assert (e["MeshSize"], e["QZoom"], e["AngularAveraged"], e["DoubleMap"], e["Interp1"], e["Interp2"]) == \
        (255, 1.5, 1, 0, 1, 1)
    assert (e["PhotonOffset"], e["PhotonStep"]) == (0.5, 2.5)
    assert (e["RestrictStackToBoundaries"], e["LowerBoundary"], e["UpperBoundary"]) == (1, 5, 500)
    assert (e["UsePixelMask_as_Flatfield"], e["FractionalCq"], e["SizeOfCqFraction"], e["InvertRotMatrix"]) == \
        (1, 0, 20, 1)
    assert e["EchoLevel"] == 0 and e["echo"] == 0 and e["Out_Report"] == "rep.txt"
    # an empty settings tree: the reference's defaults (MainRunModes.cpp:39-84)
    f = tmp_path / "empty.xml"
    f.write_text('<?xml version="1.0" encoding="utf-8"?>\n<root><EvalSettings/></root>\n')
    d = reader("eval", f)[0]
    assert (d["MeshSize"], d["QZoom"], d["DoubleMap"], d["Interp1"], d["Interp2"]) == (201, 1.0, 1, 0, 0)
    assert (d["PhotonOffset"], d["PhotonStep"], d["LowerBoundary"], d["UpperBoundary"]) == (0.0, 1.0, 0, 10)
    assert (d["SizeOfCqFraction"], d["EchoLevel"], d["XML_Path"], d["DetectorSize"]) == (10, 3, "", [0, 0])
    # a value that does not convert falls back to the default, as boost's get with a default
    f.write_text('<root><EvalSettings><OutputMesh><MeshSize>big</MeshSize></OutputMesh></EvalSettings></root>')
    assert reader("eval", f)[0]["MeshSize"] == 201


def test_example_evaluation_settings_round_trip(reader, tmp_path):
    """iac_eval -eec writes the example file (MainRunModes.cpp:354-500); it reads back."""
    f = tmp_path / "example.xml"
    subprocess.run([str(PKG / "iac_eval"), "-eec", str(f)], check=True, capture_output=True, timeout=60)
    d = reader("eval", f)[0]
    assert (d["XML_Path"], d["PixelMap_Path"], d["MeshSize"], d["DetectorSize"]) == \
        ("EventList.xml", "Pixelmap.bin", 201, [1024, 1024])
    assert d["Out_Report"] == "Report.txt" and d["EchoLevel"] == 3
