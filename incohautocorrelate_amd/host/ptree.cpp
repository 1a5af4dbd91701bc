// ptree.cpp -- the property tree of ptree.hpp and its XML, JSON and INFO readers / writers.
#include "ptree.hpp"

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace iac_host {

namespace {

std::string trim(const std::string &s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) a++;
    while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
    return s.substr(a, b - a);
}

std::vector<std::string> split_path(const std::string &path)
{
    std::vector<std::string> parts;
    size_t p = 0;
    while (true) {
        const size_t q = path.find('.', p);
        parts.push_back(path.substr(p, q == std::string::npos ? std::string::npos : q - p));
        if (q == std::string::npos) break;
        p = q + 1;
    }
    return parts;
}

[[noreturn]] void parse_error(const char *format, size_t pos, const std::string &what)
{
    throw std::runtime_error(std::string(format) + " parse error at offset " + std::to_string(pos) + ": " + what);
}

void append_utf8(std::string &out, unsigned long cp)
{
    if (cp < 0x80) {
        out += (char)cp;
    } else if (cp < 0x800) {
        out += (char)(0xC0 | (cp >> 6));
        out += (char)(0x80 | (cp & 0x3F));
    } else if (cp < 0x10000) {
        out += (char)(0xE0 | (cp >> 12));
        out += (char)(0x80 | ((cp >> 6) & 0x3F));
        out += (char)(0x80 | (cp & 0x3F));
    } else {
        out += (char)(0xF0 | (cp >> 18));
        out += (char)(0x80 | ((cp >> 12) & 0x3F));
        out += (char)(0x80 | ((cp >> 6) & 0x3F));
        out += (char)(0x80 | (cp & 0x3F));
    }
}

// ---- XML ----------------------------------------------------------------------------

struct XmlReader {
    const std::string &t;
    size_t i = 0;
    explicit XmlReader(const std::string &text) : t(text) {}

    bool starts(const char *s) const { return t.compare(i, std::strlen(s), s) == 0; }
    void skip_past(const char *s)
    {
        const size_t q = t.find(s, i);
        if (q == std::string::npos) parse_error("xml", i, std::string("missing '") + s + "'");
        i = q + std::strlen(s);
    }
    std::string name()
    {
        const size_t a = i;
        while (i < t.size() && !std::isspace((unsigned char)t[i]) && t[i] != '>' && t[i] != '/' && t[i] != '=')
            i++;
        if (i == a) parse_error("xml", i, "expected a name");
        return t.substr(a, i - a);
    }
    void skip_ws()
    {
        while (i < t.size() && std::isspace((unsigned char)t[i])) i++;
    }
    std::string decode(const std::string &s) const
    {
        std::string out;
        for (size_t k = 0; k < s.size(); k++) {
            if (s[k] != '&') {
                out += s[k];
                continue;
            }
            const size_t e = s.find(';', k);
            if (e == std::string::npos) parse_error("xml", i, "unterminated entity");
            const std::string ent = s.substr(k + 1, e - k - 1);
            if (ent == "lt") out += '<';
            else if (ent == "gt") out += '>';
            else if (ent == "amp") out += '&';
            else if (ent == "quot") out += '"';
            else if (ent == "apos") out += '\'';
            else if (!ent.empty() && ent[0] == '#')
                append_utf8(out, ent.size() > 1 && (ent[1] == 'x' || ent[1] == 'X')
                                     ? std::strtoul(ent.c_str() + 2, nullptr, 16)
                                     : std::strtoul(ent.c_str() + 1, nullptr, 10));
            else parse_error("xml", i, "unknown entity &" + ent + ";");
            k = e;
        }
        return out;
    }
    // skips comments, processing instructions and declarations; false at '<' of a tag
    bool skip_misc()
    {
        if (starts("<?")) skip_past("?>");
        else if (starts("<!--")) skip_past("-->");
        else if (starts("<!") && !starts("<![CDATA[")) skip_past(">");
        else return false;
        return true;
    }
    // one element starting at '<name'; attributes are read past and not kept
    void element(PTree &parent)
    {
        i++;  // '<'
        const std::string key = name();
        PTree node;
        while (true) {
            skip_ws();
            if (i >= t.size()) parse_error("xml", i, "unterminated tag <" + key + ">");
            if (t[i] == '/') {
                if (i + 1 >= t.size() || t[i + 1] != '>') parse_error("xml", i, "bad empty tag");
                i += 2;
                parent.children.emplace_back(key, std::move(node));
                return;
            }
            if (t[i] == '>') {
                i++;
                break;
            }
            name();
            skip_ws();
            if (i < t.size() && t[i] == '=') {
                i++;
                skip_ws();
                if (i >= t.size() || (t[i] != '"' && t[i] != '\'')) parse_error("xml", i, "attribute value");
                const char q = t[i++];
                const size_t e = t.find(q, i);
                if (e == std::string::npos) parse_error("xml", i, "unterminated attribute");
                i = e + 1;
            }
        }
        std::string text;
        while (true) {
            if (i >= t.size()) parse_error("xml", i, "missing </" + key + ">");
            if (t[i] != '<') {
                const size_t e = t.find('<', i);
                text += decode(t.substr(i, (e == std::string::npos ? t.size() : e) - i));
                i = e == std::string::npos ? t.size() : e;
                continue;
            }
            if (starts("<![CDATA[")) {
                i += 9;
                const size_t e = t.find("]]>", i);
                if (e == std::string::npos) parse_error("xml", i, "unterminated CDATA");
                text += t.substr(i, e - i);
                i = e + 3;
                continue;
            }
            if (skip_misc()) continue;
            if (starts("</")) {
                i += 2;
                const std::string end = name();
                if (end != key) parse_error("xml", i, "</" + end + "> closes <" + key + ">");
                skip_ws();
                if (i >= t.size() || t[i] != '>') parse_error("xml", i, "bad end tag");
                i++;
                break;
            }
            element(node);
        }
        // leaf values without their surrounding blanks (indented files); element text of
        // inner nodes is the blanks between children and is dropped
        node.data = node.children.empty() ? trim(text) : std::string();
        parent.children.emplace_back(key, std::move(node));
    }
    PTree document()
    {
        PTree root;
        while (true) {
            skip_ws();
            if (i >= t.size()) break;
            if (skip_misc()) continue;
            if (t[i] != '<') parse_error("xml", i, "text outside the root element");
            element(root);
        }
        return root;
    }
};

std::string xml_escape(const std::string &s)
{
    std::string out;
    for (char c : s) {
        switch (c) {
        case '<': out += "&lt;"; break;
        case '>': out += "&gt;"; break;
        case '&': out += "&amp;"; break;
        case '"': out += "&quot;"; break;
        case '\'': out += "&apos;"; break;
        default: out += c;
        }
    }
    return out;
}

void xml_node(std::string &out, const std::string &key, const PTree &n)
{
    out += "<" + key + ">" + xml_escape(n.data);
    for (const auto &c : n.children) xml_node(out, c.first, c.second);
    out += "</" + key + ">";
}

// ---- JSON ---------------------------------------------------------------------------

struct JsonReader {
    const std::string &t;
    size_t i = 0;
    explicit JsonReader(const std::string &text) : t(text) {}
    void ws()
    {
        while (i < t.size() && std::isspace((unsigned char)t[i])) i++;
    }
    char peek()
    {
        ws();
        if (i >= t.size()) parse_error("json", i, "unexpected end");
        return t[i];
    }
    void expect(char c)
    {
        if (peek() != c) parse_error("json", i, std::string("expected '") + c + "'");
        i++;
    }
    std::string str()
    {
        expect('"');
        std::string out;
        while (true) {
            if (i >= t.size()) parse_error("json", i, "unterminated string");
            const char c = t[i++];
            if (c == '"') break;
            if (c != '\\') {
                out += c;
                continue;
            }
            if (i >= t.size()) parse_error("json", i, "bad escape");
            const char e = t[i++];
            switch (e) {
            case '"': out += '"'; break;
            case '\\': out += '\\'; break;
            case '/': out += '/'; break;
            case 'b': out += '\b'; break;
            case 'f': out += '\f'; break;
            case 'n': out += '\n'; break;
            case 'r': out += '\r'; break;
            case 't': out += '\t'; break;
            case 'u': {
                if (i + 4 > t.size()) parse_error("json", i, "bad \\u escape");
                append_utf8(out, std::strtoul(t.substr(i, 4).c_str(), nullptr, 16));
                i += 4;
                break;
            }
            default: parse_error("json", i, "bad escape");
            }
        }
        return out;
    }
    void value(PTree &n)
    {
        const char c = peek();
        if (c == '{') {
            i++;
            if (peek() == '}') {
                i++;
                return;
            }
            while (true) {
                const std::string k = str();
                expect(':');
                PTree child;
                value(child);
                n.children.emplace_back(k, std::move(child));
                if (peek() == ',') {
                    i++;
                    continue;
                }
                expect('}');
                return;
            }
        }
        if (c == '[') {
            i++;
            if (peek() == ']') {
                i++;
                return;
            }
            while (true) {
                PTree child;
                value(child);
                n.children.emplace_back("", std::move(child));
                if (peek() == ',') {
                    i++;
                    continue;
                }
                expect(']');
                return;
            }
        }
        if (c == '"') {
            n.data = str();
            return;
        }
        const size_t a = i;
        while (i < t.size() && !std::isspace((unsigned char)t[i]) && t[i] != ',' && t[i] != '}' && t[i] != ']') i++;
        if (i == a) parse_error("json", i, "expected a value");
        n.data = t.substr(a, i - a);
    }
};

std::string json_escape(const std::string &s)
{
    std::string out;
    for (unsigned char c : s) {
        if (c == '"') out += "\\\"";
        else if (c == '\\') out += "\\\\";
        else if (c == '\n') out += "\\n";
        else if (c == '\t') out += "\\t";
        else if (c == '\r') out += "\\r";
        else if (c < 0x20) {
            char b[8];
            std::snprintf(b, sizeof b, "\\u%04x", c);
            out += b;
        } else out += (char)c;
    }
    return out;
}

void json_node(std::string &out, const PTree &n, int depth)
{
    const std::string pad((size_t)depth * 4, ' ');
    if (n.children.empty()) {
        out += "\"" + json_escape(n.data) + "\"";
        return;
    }
    const bool array = std::all_of(n.children.begin(), n.children.end(), [](const auto &c) { return c.first.empty(); });
    out += array ? "[\n" : "{\n";
    for (size_t k = 0; k < n.children.size(); k++) {
        out += pad + "    ";
        if (!array) out += "\"" + json_escape(n.children[k].first) + "\": ";
        json_node(out, n.children[k].second, depth + 1);
        out += k + 1 < n.children.size() ? ",\n" : "\n";
    }
    out += pad + (array ? "]" : "}");
}

// ---- INFO ---------------------------------------------------------------------------

struct InfoToken {
    enum Kind { Word, Open, Close, Newline, End } kind;
    std::string text;
    size_t pos;
};

std::vector<InfoToken> info_tokens(const std::string &t)
{
    std::vector<InfoToken> out;
    size_t i = 0;
    while (i < t.size()) {
        const char c = t[i];
        if (c == '\n') {
            out.push_back({InfoToken::Newline, "", i++});
        } else if (std::isspace((unsigned char)c)) {
            i++;
        } else if (c == ';') {
            while (i < t.size() && t[i] != '\n') i++;
        } else if (c == '{') {
            out.push_back({InfoToken::Open, "", i++});
        } else if (c == '}') {
            out.push_back({InfoToken::Close, "", i++});
        } else if (c == '"') {
            const size_t a = i++;
            std::string s;
            while (true) {
                if (i >= t.size() || t[i] == '\n') parse_error("info", a, "unterminated string");
                const char d = t[i++];
                if (d == '"') break;
                if (d != '\\') {
                    s += d;
                    continue;
                }
                if (i >= t.size()) parse_error("info", i, "bad escape");
                const char e = t[i++];
                switch (e) {
                case '0': s += '\0'; break;
                case 'a': s += '\a'; break;
                case 'b': s += '\b'; break;
                case 'f': s += '\f'; break;
                case 'n': s += '\n'; break;
                case 'r': s += '\r'; break;
                case 't': s += '\t'; break;
                case 'v': s += '\v'; break;
                default: s += e; break;  // \" \' \\ and the rest as themselves
                }
            }
            out.push_back({InfoToken::Word, s, a});
        } else {
            const size_t a = i;
            while (i < t.size() && !std::isspace((unsigned char)t[i]) && t[i] != ';' && t[i] != '{' && t[i] != '}')
                i++;
            out.push_back({InfoToken::Word, t.substr(a, i - a), a});
        }
    }
    out.push_back({InfoToken::End, "", t.size()});
    return out;
}

struct InfoReader {
    std::vector<InfoToken> tok;
    size_t k = 0;
    void skip_newlines()
    {
        while (tok[k].kind == InfoToken::Newline) k++;
    }
    // items up to a '}' (inner) or the end (top level)
    void items(PTree &n, bool inner)
    {
        while (true) {
            skip_newlines();
            const InfoToken &t = tok[k];
            if (t.kind == InfoToken::End) {
                if (inner) parse_error("info", t.pos, "missing '}'");
                return;
            }
            if (t.kind == InfoToken::Close) {
                if (!inner) parse_error("info", t.pos, "unmatched '}'");
                k++;
                return;
            }
            if (t.kind != InfoToken::Word) parse_error("info", t.pos, "expected a key");
            PTree child;
            const std::string key = t.text;
            k++;
            if (tok[k].kind == InfoToken::Word) child.data = tok[k++].text;
            size_t save = k;
            skip_newlines();
            if (tok[k].kind == InfoToken::Open) {
                k++;
                items(child, true);
            } else {
                k = save;
            }
            n.children.emplace_back(key, std::move(child));
        }
    }
};

std::string info_value(const std::string &s)
{
    bool plain = !s.empty();
    for (unsigned char c : s)
        if (std::isspace(c) || c == ';' || c == '{' || c == '}' || c == '"' || c == '\\') plain = false;
    if (plain) return s;
    std::string out = "\"";
    for (char c : s) {
        if (c == '"' || c == '\\') out += '\\';
        if (c == '\n') {
            out += "\\n";
            continue;
        }
        out += c;
    }
    return out + "\"";
}

void info_node(std::string &out, const std::string &key, const PTree &n, int depth)
{
    const std::string pad((size_t)depth * 4, ' ');
    out += pad + info_value(key);
    if (!n.data.empty() || n.children.empty()) out += " " + info_value(n.data);
    out += "\n";
    if (!n.children.empty()) {
        out += pad + "{\n";
        for (const auto &c : n.children) info_node(out, c.first, c.second, depth + 1);
        out += pad + "}\n";
    }
}

std::string read_file(const std::string &filename)
{
    std::ifstream f(filename, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + filename);
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

std::string extension(const std::string &filename)
{
    const size_t d = filename.rfind('.');
    std::string e = d == std::string::npos ? filename : filename.substr(d + 1);
    std::transform(e.begin(), e.end(), e.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    return e;
}

}  // namespace

// ---- tree access --------------------------------------------------------------------

const PTree *PTree::child(const std::string &key) const
{
    for (const auto &c : children)
        if (c.first == key) return &c.second;
    return nullptr;
}

PTree *PTree::child(const std::string &key)
{
    for (auto &c : children)
        if (c.first == key) return &c.second;
    return nullptr;
}

const PTree *PTree::find(const std::string &path) const
{
    const PTree *n = this;
    for (const std::string &p : split_path(path)) {
        n = n->child(p);
        if (!n) return nullptr;
    }
    return n;
}

PTree &PTree::put(const std::string &path, const std::string &value)
{
    PTree *n = this;
    for (const std::string &p : split_path(path)) {
        PTree *c = n->child(p);
        if (!c) {
            n->children.emplace_back(p, PTree());
            c = &n->children.back().second;
        }
        n = c;
    }
    n->data = value;
    return *n;
}

bool parse_value(const std::string &s, long long &out)
{
    const std::string v = trim(s);
    if (v.empty()) return false;
    errno = 0;
    char *end = nullptr;
    const long long x = std::strtoll(v.c_str(), &end, 10);
    if (errno || *end) return false;
    out = x;
    return true;
}

bool parse_value(const std::string &s, double &out)
{
    const std::string v = trim(s);
    if (v.empty()) return false;
    char *end = nullptr;
    const double x = std::strtod(v.c_str(), &end);
    if (*end) return false;
    out = x;
    return true;
}

bool parse_value(const std::string &s, bool &out)
{
    const std::string v = trim(s);
    if (v == "1" || v == "true") out = true;
    else if (v == "0" || v == "false") out = false;
    else return false;
    return true;
}

std::string format_float(float v)
{
    char b[32];
    std::snprintf(b, sizeof b, "%.9g", (double)v);
    return b;
}

std::string PTree::get(const std::string &path, const std::string &dflt) const
{
    const PTree *n = find(path);
    return n ? n->data : dflt;
}

int PTree::get(const std::string &path, int dflt) const
{
    const PTree *n = find(path);
    long long v = 0;
    return n && parse_value(n->data, v) && v >= INT32_MIN && v <= INT32_MAX ? (int)v : dflt;
}

unsigned int PTree::get(const std::string &path, unsigned int dflt) const
{
    const PTree *n = find(path);
    long long v = 0;
    return n && parse_value(n->data, v) && v >= 0 && v <= (long long)UINT32_MAX ? (unsigned int)v : dflt;
}

float PTree::get(const std::string &path, float dflt) const
{
    const PTree *n = find(path);
    double v = 0;
    return n && parse_value(n->data, v) ? (float)v : dflt;
}

double PTree::get(const std::string &path, double dflt) const
{
    const PTree *n = find(path);
    double v = 0;
    return n && parse_value(n->data, v) ? v : dflt;
}

bool PTree::get(const std::string &path, bool dflt) const
{
    const PTree *n = find(path);
    bool v = false;
    return n && parse_value(n->data, v) ? v : dflt;
}

std::string PTree::get_required(const std::string &path) const
{
    const PTree *n = find(path);
    if (!n) throw std::runtime_error("No such node (" + path + ")");
    return n->data;
}

int PTree::get_required_int(const std::string &path) const
{
    long long v = 0;
    if (!parse_value(get_required(path), v) || v < INT32_MIN || v > INT32_MAX)
        throw std::runtime_error("conversion of data to type \"int\" failed (" + path + ")");
    return (int)v;
}

float PTree::get_required_float(const std::string &path) const
{
    double v = 0;
    if (!parse_value(get_required(path), v))
        throw std::runtime_error("conversion of data to type \"float\" failed (" + path + ")");
    return (float)v;
}

// ---- formats ------------------------------------------------------------------------

PTree parse_xml(const std::string &text) { return XmlReader(text).document(); }

PTree parse_json(const std::string &text)
{
    JsonReader r(text);
    PTree root;
    r.value(root);
    r.ws();
    if (r.i != text.size()) parse_error("json", r.i, "text after the document");
    return root;
}

PTree parse_info(const std::string &text)
{
    InfoReader r;
    r.tok = info_tokens(text);
    PTree root;
    r.items(root, false);
    return root;
}

std::string write_xml(const PTree &pt)
{
    std::string out = "<?xml version=\"1.0\" encoding=\"utf-8\"?>\n";
    for (const auto &c : pt.children) xml_node(out, c.first, c.second);
    return out + "\n";
}

std::string write_json(const PTree &pt)
{
    std::string out;
    json_node(out, pt, 0);
    return out + "\n";
}

std::string write_info(const PTree &pt)
{
    std::string out;
    for (const auto &c : pt.children) info_node(out, c.first, c.second, 0);
    return out;
}

PTree load_property_tree(const std::string &filename)
{
    const std::string text = read_file(filename), ext = extension(filename);
    if (ext == "json") return parse_json(text);
    if (ext == "info") return parse_info(text);
    return parse_xml(text);
}

void save_property_tree(const PTree &pt, const std::string &filename)
{
    const std::string ext = extension(filename);
    std::string text;
    if (ext == "json") text = write_json(pt);
    else if (ext == "info") text = write_info(pt);
    else text = write_xml(pt);
    std::ofstream f(filename, std::ios::binary);
    if (!f) throw std::runtime_error("cannot write " + filename);
    f << text;
    if (!f) throw std::runtime_error("short write to " + filename);
}

}  // namespace iac_host
