// Minimal JSON DOM for the solver's setup files (vehicle, tire, suspension parameters).
// Numbers are converted with strtod (correctly rounded); the reference's parser rounds the
// same way for every number with at most 15 significant digits, which covers all bundled files.
#pragma once
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace lfdm {

struct JValue {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<JValue> arr;
    std::vector<std::pair<std::string, JValue>> obj;

    const JValue* find(const std::string& key) const;
    bool has(const std::string& key) const { return find(key) != nullptr; }
};

// Parses `text`; on failure returns false and sets `err` to a message with the byte offset.
bool json_parse(const std::string& text, JValue& out, std::string& err);
// Reads and parses a file; `err` names the file on failure.
bool json_load_file(const std::string& path, JValue& out, std::string& err);

}  // namespace lfdm
