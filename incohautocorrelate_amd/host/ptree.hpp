// ptree.hpp -- a small property tree with the XML, JSON and INFO readers and writers the
// reference's configuration files need, in place of boost::property_tree (IAC/Settings.h:44-84
// picks the format by file extension: "json", "info", else XML).  Paths are dot-separated
// and case-sensitive; values are strings converted on access the way boost's stream
// translator converts them (the whole string must parse, surrounding blanks allowed; bools
// accept 0 / 1 and true / false).
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace iac_host {

struct PTree {
    std::string data;
    std::vector<std::pair<std::string, PTree>> children;

    // the first child named `key`, or nullptr
    const PTree *child(const std::string &key) const;
    PTree *child(const std::string &key);
    // the node at a dot-separated path, or nullptr
    const PTree *find(const std::string &path) const;
    // sets the node at `path`, creating the missing ones (boost ptree::put)
    PTree &put(const std::string &path, const std::string &value);

    // boost ptree::get<T>(path, default): the default when the path is missing or its value
    // does not convert; get_required throws std::runtime_error instead
    std::string get(const std::string &path, const std::string &dflt) const;
    int get(const std::string &path, int dflt) const;
    unsigned int get(const std::string &path, unsigned int dflt) const;
    float get(const std::string &path, float dflt) const;
    double get(const std::string &path, double dflt) const;
    bool get(const std::string &path, bool dflt) const;
    std::string get_required(const std::string &path) const;
    int get_required_int(const std::string &path) const;
    float get_required_float(const std::string &path) const;
};

// text conversions of a value (false when the text does not convert)
bool parse_value(const std::string &s, long long &out);
bool parse_value(const std::string &s, double &out);
bool parse_value(const std::string &s, bool &out);
// a float as boost writes it: the shortest of max_digits10 significant digits
std::string format_float(float v);

PTree parse_xml(const std::string &text);
PTree parse_json(const std::string &text);
PTree parse_info(const std::string &text);
std::string write_xml(const PTree &pt);
std::string write_json(const PTree &pt);
std::string write_info(const PTree &pt);

// Settings::LoadPropertyTree / SavePropertyTree (IAC/Settings.h:44-84): the format from the
// file name's extension (case-insensitive) -- json, info, otherwise xml.
PTree load_property_tree(const std::string &filename);
void save_property_tree(const PTree &pt, const std::string &filename);

}  // namespace iac_host
