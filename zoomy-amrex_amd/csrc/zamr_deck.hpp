#pragma once
// ParmParse-style "inputs" deck reader for the keys the reference driver
// queries (Source/ZoomyAmr.cpp:29-83) and build.py:write_inputs writes
// (build.py:55-120): `prefix.key = v1 v2 ...`, `#` comments, optional quotes.
#include <cctype>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace zamr {

class Deck {
   public:
    explicit Deck(const std::string& text) {
        std::istringstream in(text);
        std::string line;
        while (std::getline(in, line)) {
            const auto hash = line.find('#');
            if (hash != std::string::npos) line = line.substr(0, hash);
            const auto eq = line.find('=');
            if (eq == std::string::npos) continue;
            std::string key = trim(line.substr(0, eq));
            if (key.empty()) continue;
            std::vector<std::string> vals;
            std::istringstream vs(line.substr(eq + 1));
            std::string tok;
            while (vs >> tok) {
                if (tok.size() >= 2 && (tok.front() == '"' || tok.front() == '\'') && tok.back() == tok.front())
                    tok = tok.substr(1, tok.size() - 2);
                vals.push_back(tok);
            }
            kv_[key] = vals;  // last definition wins, as ParmParse::query does
        }
    }

    bool has(const std::string& k) const { return kv_.count(k) != 0; }
    const std::vector<std::string>& vals(const std::string& k) const { return kv_.at(k); }

    bool query(const std::string& k, std::string& v) const {
        if (!has(k) || kv_.at(k).empty()) return false;
        v = kv_.at(k)[0];
        return true;
    }
    bool query(const std::string& k, double& v) const {
        std::string s;
        if (!query(k, s)) return false;
        v = to_double(k, s);
        return true;
    }
    bool query(const std::string& k, int& v) const {
        std::string s;
        if (!query(k, s)) return false;
        v = (int)to_double(k, s);
        return true;
    }
    bool query(const std::string& k, bool& v) const {
        std::string s;
        if (!query(k, s)) return false;
        std::string l;
        for (char c : s) l += (char)std::tolower((unsigned char)c);
        if (l == "true" || l == "1" || l == "t" || l == "yes") v = true;
        else if (l == "false" || l == "0" || l == "f" || l == "no") v = false;
        else throw std::invalid_argument("deck key " + k + ": not a boolean: " + s);
        return true;
    }
    bool queryarr(const std::string& k, std::vector<double>& v) const {
        if (!has(k)) return false;
        v.clear();
        for (auto& s : kv_.at(k)) v.push_back(to_double(k, s));
        return true;
    }
    bool queryarr(const std::string& k, std::vector<std::string>& v) const {
        if (!has(k)) return false;
        v = kv_.at(k);
        return true;
    }

   private:
    static std::string trim(const std::string& s) {
        size_t a = 0, b = s.size();
        while (a < b && std::isspace((unsigned char)s[a])) ++a;
        while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
        return s.substr(a, b - a);
    }
    static double to_double(const std::string& k, const std::string& s) {
        try {
            size_t pos = 0;
            double d = std::stod(s, &pos);
            if (pos != s.size()) throw std::invalid_argument(s);
            return d;
        } catch (...) {
            throw std::invalid_argument("deck key " + k + ": not a number: " + s);
        }
    }
    std::map<std::string, std::vector<std::string>> kv_;
};

}  // namespace zamr
