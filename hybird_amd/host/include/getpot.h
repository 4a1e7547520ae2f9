// Command-line and configuration-file reader with the interface that hybird.cpp
// and the parameter macros use (the reference vendors the GetPot library,
// getpot.h; this is an independent implementation of the subset it needs).
//
// File syntax: one "name = value" per line; '#' starts a comment; a value may
// be quoted with single or double quotes; "[section]" lines prefix the names
// that follow with "section/". The command line is a list of words; search()
// positions a cursor on a flag and next() reads the word after it.
#ifndef HYBIRD_AMD_GETPOT_H
#define HYBIRD_AMD_GETPOT_H

#include <cstdlib>
#include <fstream>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <vector>

class GetPot {
public:
    GetPot() = default;

    GetPot(int argc, char** argv) {
        for (int i = 1; i < argc; ++i) args_.push_back(argv[i] ? argv[i] : "");
        used_.assign(args_.size(), false);
    }

    explicit GetPot(const std::string& file) { read_file(file); }
    explicit GetPot(const char* file) { read_file(file ? file : ""); }

    // ---- command line
    bool search(const char* flag) {
        for (size_t i = 0; i < args_.size(); ++i) {
            if (args_[i] == flag) {
                cursor_ = (long)i;
                used_[i] = true;
                return true;
            }
        }
        return false;
    }

    int next(int def) { return to_int(next_word(), def); }
    unsigned next(unsigned def) { return (unsigned)to_int(next_word(), (int)def); }
    bool next(bool def) { return to_int(next_word(), def ? 1 : 0) != 0; }
    double next(double def) { return to_double(next_word(), def); }
    std::string next(const std::string& def) {
        const std::string* w = next_word();
        return w ? *w : def;
    }
    std::string next(const char* def) { return next(std::string(def ? def : "")); }

    std::vector<std::string> unidentified_arguments() const {
        std::vector<std::string> out;
        for (size_t i = 0; i < args_.size(); ++i)
            if (!used_[i] && !args_[i].empty() && args_[i][0] == '-') out.push_back(args_[i]);
        return out;
    }

    // ---- configuration file
    bool have_variable(const char* name) const { return vars_.count(name) != 0; }

    int operator()(const char* name, int def) { return to_int(value(name), def); }
    unsigned operator()(const char* name, unsigned def) { return (unsigned)to_int(value(name), (int)def); }
    bool operator()(const char* name, bool def) { return to_int(value(name), def ? 1 : 0) != 0; }
    double operator()(const char* name, double def) { return to_double(value(name), def); }
    std::string operator()(const char* name, const std::string& def) {
        const std::string* v = value(name);
        return v ? *v : def;
    }
    std::string operator()(const char* name, const char* def) { return (*this)(name, std::string(def ? def : "")); }

    std::vector<std::string> unidentified_variables() const {
        std::vector<std::string> out;
        for (const auto& kv : vars_)
            if (!requested_.count(kv.first)) out.push_back(kv.first);
        return out;
    }

private:
    std::vector<std::string> args_;
    std::vector<bool> used_;
    long cursor_ = -1;
    std::map<std::string, std::string> vars_;
    std::set<std::string> requested_;

    const std::string* next_word() {
        if (cursor_ < 0 || cursor_ + 1 >= (long)args_.size()) return nullptr;
        ++cursor_;
        used_[cursor_] = true;
        return &args_[cursor_];
    }

    const std::string* value(const char* name) {
        requested_.insert(name);
        auto it = vars_.find(name);
        return it == vars_.end() ? nullptr : &it->second;
    }

    // Integer conversion: integral text only; anything else keeps the default.
    static int to_int(const std::string* s, int def) {
        if (!s) return def;
        char* end = nullptr;
        const long v = std::strtol(s->c_str(), &end, 10);
        if (end == s->c_str() || *end != '\0') {
            // a floating-point literal for an integer parameter truncates
            const double d = std::strtod(s->c_str(), &end);
            if (end == s->c_str() || *end != '\0') return def;
            return (int)d;
        }
        return (int)v;
    }
    static double to_double(const std::string* s, double def) {
        if (!s) return def;
        char* end = nullptr;
        const double v = std::strtod(s->c_str(), &end);
        if (end == s->c_str() || *end != '\0') return def;
        return v;
    }

    static std::string trim(const std::string& s) {
        const size_t a = s.find_first_not_of(" \t\r\n");
        if (a == std::string::npos) return "";
        const size_t b = s.find_last_not_of(" \t\r\n");
        return s.substr(a, b - a + 1);
    }

    void read_file(const std::string& file) {
        std::ifstream in(file.c_str());
        std::string line, section;
        while (std::getline(in, line)) {
            // strip comments outside quotes
            std::string clean;
            char quote = 0;
            for (char ch : line) {
                if (quote) {
                    if (ch == quote) quote = 0;
                } else if (ch == '\'' || ch == '"') {
                    quote = ch;
                } else if (ch == '#') {
                    break;
                }
                clean.push_back(ch);
            }
            clean = trim(clean);
            if (clean.empty()) continue;
            if (clean.front() == '[' && clean.back() == ']') {
                section = trim(clean.substr(1, clean.size() - 2));
                if (!section.empty() && section != "/") section += "/";
                else section.clear();
                continue;
            }
            const size_t eq = clean.find('=');
            if (eq == std::string::npos) continue;
            std::string name = trim(clean.substr(0, eq));
            std::string val = trim(clean.substr(eq + 1));
            if (val.size() >= 2 && (val.front() == '\'' || val.front() == '"') && val.back() == val.front())
                val = val.substr(1, val.size() - 2);
            if (!name.empty()) vars_[section + name] = val;
        }
    }
};

#endif  // HYBIRD_AMD_GETPOT_H
