// Minimal reader for the per-algorithm argument files (--args_path): one flat JSON object
// of numbers, strings and booleans, read with defaults like nlohmann::json::value(key, default)
// in the reference wrappers (e.g. src/one_to_one.hpp:185, src/one_to_many.hpp:362-365).
#pragma once

#include <cctype>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>

namespace xcli {

class json_args {
public:
    json_args() = default;

    static json_args parse(const std::string& text) {
        json_args a;
        size_t i = 0;
        auto ws = [&] {
            while (i < text.size() && std::isspace((unsigned char)text[i])) ++i;
        };
        auto str = [&]() -> std::string {
            if (text[i] != '"') throw std::runtime_error("JSON: expected string");
            ++i;
            std::string s;
            while (i < text.size() && text[i] != '"') {
                if (text[i] == '\\' && i + 1 < text.size()) ++i;
                s += text[i++];
            }
            if (i >= text.size()) throw std::runtime_error("JSON: unterminated string");
            ++i;
            return s;
        };
        ws();
        if (i >= text.size() || text[i] != '{') throw std::runtime_error("JSON: arguments must be an object");
        ++i;
        ws();
        if (i < text.size() && text[i] == '}') return a;
        while (true) {
            ws();
            const std::string key = str();
            ws();
            if (i >= text.size() || text[i] != ':') throw std::runtime_error("JSON: expected ':'");
            ++i;
            ws();
            if (i >= text.size()) throw std::runtime_error("JSON: missing value");
            std::string val;
            if (text[i] == '"') {
                val = "s" + str();
            } else if (text[i] == '{' || text[i] == '[') {
                throw std::runtime_error("JSON: nested values are not supported for key " + key);
            } else {
                size_t j = i;
                while (j < text.size() && text[j] != ',' && text[j] != '}' && !std::isspace((unsigned char)text[j])) ++j;
                val = "v" + text.substr(i, j - i);
                i = j;
            }
            a.values_[key] = val;
            ws();
            if (i < text.size() && text[i] == ',') {
                ++i;
                continue;
            }
            if (i < text.size() && text[i] == '}') break;
            throw std::runtime_error("JSON: expected ',' or '}'");
        }
        return a;
    }

    static json_args load(const std::string& path) {
        std::ifstream in(path);
        if (!in) throw std::runtime_error("Could not open args file " + path);
        std::stringstream ss;
        ss << in.rdbuf();
        return parse(ss.str());
    }

    long get_int(const std::string& key, long def) const {
        auto it = values_.find(key);
        if (it == values_.end()) return def;
        const std::string& v = it->second;
        if (v[0] != 'v') throw std::runtime_error("argument " + key + " must be a number");
        if (v == "vtrue") return 1;
        if (v == "vfalse") return 0;
        return std::stol(v.substr(1));
    }

    bool get_bool(const std::string& key, bool def) const {
        auto it = values_.find(key);
        if (it == values_.end()) return def;
        if (it->second == "vtrue") return true;
        if (it->second == "vfalse") return false;
        return get_int(key, def) != 0;
    }

    std::string get_string(const std::string& key, const std::string& def) const {
        auto it = values_.find(key);
        if (it == values_.end()) return def;
        return it->second.substr(1);
    }

private:
    std::map<std::string, std::string> values_;
};

}  // namespace xcli
