// param_in.h -- param.in as the reference reads it (parameters::map_input and
// get/getString, src_v4.5.7/parameters.cpp:530-620): "key = value" lines,
// lines starting with '#' skipped, a trailing '\' joins the next line, values
// converted with atof and multiplied by the unit.  Keys are taken as the text
// before '=' trimmed (the reference also drops the character just before '=',
// which is the blank of "key = value"; both read such files alike).
#pragma once
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <map>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace dmdrun {

class ParamIn {
public:
    static std::string trim(std::string s)
    {
        s.erase(0, s.find_first_not_of(" \t"));
        return s.erase(s.find_last_not_of(" \n\r\t") + 1);
    }

    static ParamIn parse(std::istream &in)
    {
        ParamIn p;
        std::string last;
        bool join = false;
        for (std::string line; std::getline(in, line);) {
            if (!line.empty() && line[0] == '#') continue;
            std::string l2 = trim(line);
            const bool will_join = !l2.empty() && l2.back() == '\\';
            if (will_join) l2.back() = ' ';
            if (join) {
                p.map_[last] += " " + l2;
            } else {
                const size_t eq = l2.find('=');
                if (eq != std::string::npos) {
                    last = trim(l2.substr(0, eq));
                    p.map_[last] = trim(l2.substr(eq + 1));
                }
            }
            join = will_join;
        }
        return p;
    }

    static ParamIn read(const std::string &path)
    {
        std::ifstream f(path);
        if (!f) throw std::runtime_error("input file " + path + " does not exist");
        return parse(f);
    }

    bool has(const std::string &k) const { return map_.count(k) != 0; }

    double get(const std::string &k, double def, double unit = 1.0) const
    {
        used_.insert(k);
        auto it = map_.find(k);
        if (it == map_.end()) return def * unit;
        return std::atof(it->second.c_str()) * unit;
    }

    std::string str(const std::string &k, const std::string &def) const
    {
        used_.insert(k);
        auto it = map_.find(k);
        return it == map_.end() ? def : it->second;
    }

    // keys given in the file that this program never asked for
    std::vector<std::string> unread() const
    {
        std::vector<std::string> out;
        for (auto &kv : map_)
            if (!used_.count(kv.first)) out.push_back(kv.first);
        return out;
    }

private:
    std::map<std::string, std::string> map_;
    mutable std::set<std::string> used_;
};

}  // namespace dmdrun
