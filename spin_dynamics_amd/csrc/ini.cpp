// ini.cpp -- the .ini subset read_inp needs (spinsys.f90:60-135), in place of
// the FiNeR file_ini calls load / get / count_values / has_section / Ns.
//
// Grammar: "[section]" lines; "key = value" lines; values are lists separated
// by blanks or commas; a ';', '#' or '!' starts a comment; single or double
// quotes around a whole value are removed.  Keys and sections are matched
// after trimming, case-sensitively.
#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <utility>
#include <vector>

#include "../../include/spinchem.h"

struct spc_ini {
    struct Opt {
        std::string key, value;
    };
    struct Sec {
        std::string name;
        std::vector<Opt> opts;
    };
    std::vector<Sec> secs;

    const Opt *find(const char *section, const char *option) const
    {
        if (!section || !option) return nullptr;
        for (const auto &s : secs)
            if (s.name == section)
                for (const auto &o : s.opts)
                    if (o.key == option) return &o;
        return nullptr;
    }
};

namespace {

std::string trim(const std::string &s)
{
    size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}

std::string strip_comment(const std::string &s)
{
    char q = 0;
    for (size_t i = 0; i < s.size(); ++i) {
        char c = s[i];
        if (q) {
            if (c == q) q = 0;
        } else if (c == '\'' || c == '"') {
            q = c;
        } else if (c == ';' || c == '#' || c == '!') {
            return s.substr(0, i);
        }
    }
    return s;
}

std::string unquote(const std::string &v)
{
    if (v.size() >= 2 && (v[0] == '\'' || v[0] == '"') && v.back() == v[0]) return v.substr(1, v.size() - 2);
    return v;
}

std::vector<std::string> tokens(const std::string &v)
{
    std::vector<std::string> out;
    std::string cur;
    for (char c : v) {
        if (c == ' ' || c == '\t' || c == ',') {
            if (!cur.empty()) out.push_back(cur), cur.clear();
        } else {
            cur.push_back(c);
        }
    }
    if (!cur.empty()) out.push_back(cur);
    return out;
}

} // namespace

extern "C" {

int spc_ini_open(const char *path, spc_ini **out)
{
    if (!path || !out) return SPC_ERR_INVALID;
    *out = nullptr;
    std::ifstream f(path);
    if (!f) return SPC_ERR_INVALID;
    spc_ini *ini = new spc_ini();
    std::string line;
    while (std::getline(f, line)) {
        std::string t = trim(strip_comment(line));
        if (t.empty()) continue;
        if (t.front() == '[') {
            size_t e = t.find(']');
            spc_ini::Sec s;
            s.name = trim(t.substr(1, e == std::string::npos ? std::string::npos : e - 1));
            ini->secs.push_back(std::move(s));
            continue;
        }
        size_t eq = t.find('=');
        if (eq == std::string::npos || ini->secs.empty()) continue;
        spc_ini::Opt o;
        o.key = trim(t.substr(0, eq));
        o.value = unquote(trim(t.substr(eq + 1)));
        ini->secs.back().opts.push_back(std::move(o));
    }
    *out = ini;
    return SPC_OK;
}

void spc_ini_close(spc_ini *ini) { delete ini; }

int spc_ini_num_sections(const spc_ini *ini) { return ini ? (int)ini->secs.size() : 0; }

int spc_ini_has_section(const spc_ini *ini, const char *section)
{
    if (!ini || !section) return 0;
    for (const auto &s : ini->secs)
        if (s.name == section) return 1;
    return 0;
}

int spc_ini_count_values(const spc_ini *ini, const char *section, const char *option)
{
    const spc_ini::Opt *o = ini ? ini->find(section, option) : nullptr;
    return o ? (int)tokens(o->value).size() : -1;
}

int spc_ini_get_string(const spc_ini *ini, const char *section, const char *option, char *buf, int32_t buflen)
{
    const spc_ini::Opt *o = ini ? ini->find(section, option) : nullptr;
    if (!o || !buf || buflen < 1) return SPC_ERR_INVALID;
    std::strncpy(buf, o->value.c_str(), (size_t)buflen - 1);
    buf[buflen - 1] = '\0';
    return SPC_OK;
}

int spc_ini_get_doubles(const spc_ini *ini, const char *section, const char *option, double *out, int32_t n)
{
    const spc_ini::Opt *o = ini ? ini->find(section, option) : nullptr;
    if (!o || !out) return SPC_ERR_INVALID;
    auto tk = tokens(o->value);
    if ((int)tk.size() < n) return SPC_ERR_INVALID;
    for (int i = 0; i < n; ++i) {
        std::string s = tk[i];
        for (char &c : s)
            if (c == 'd' || c == 'D') c = 'e'; // Fortran exponent letter
        char *end = nullptr;
        errno = 0;
        out[i] = std::strtod(s.c_str(), &end);
        if (end == s.c_str() || *end != '\0' || errno) return SPC_ERR_INVALID;
    }
    return SPC_OK;
}

int spc_ini_get_int64s(const spc_ini *ini, const char *section, const char *option, int64_t *out, int32_t n)
{
    const spc_ini::Opt *o = ini ? ini->find(section, option) : nullptr;
    if (!o || !out) return SPC_ERR_INVALID;
    auto tk = tokens(o->value);
    if ((int)tk.size() < n) return SPC_ERR_INVALID;
    for (int i = 0; i < n; ++i) {
        char *end = nullptr;
        errno = 0;
        out[i] = std::strtoll(tk[i].c_str(), &end, 10);
        if (end == tk[i].c_str() || *end != '\0' || errno) return SPC_ERR_INVALID;
    }
    return SPC_OK;
}

} // extern "C"
