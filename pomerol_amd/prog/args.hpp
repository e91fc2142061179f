// Command-line options of the ED programs: "--name value", "--name=value", boolean "--flag",
// comma-separated lists, -h/--help with defaults.  The option names, defaults and help strings
// are those of the reference programs (prog/quantum_model.cpp:28-42, anderson_model.hpp:35-38,
// hubbard2d_model.hpp:37-42).
#pragma once

#include <algorithm>
#include <cstdlib>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace progargs {

class Parser {
    struct Option {
        std::string name, help, def;
        bool flag;
        std::string value;
        bool given = false;
    };
    std::string description;
    std::vector<Option> options;

    Option& find(std::string const& name) {
        for(auto& o : options)
            if(o.name == name)
                return o;
        throw std::logic_error("unknown option --" + name);
    }
    Option const& find(std::string const& name) const { return const_cast<Parser*>(this)->find(name); }

    std::string text(std::string const& name) const {
        Option const& o = find(name);
        return o.given ? o.value : o.def;
    }

public:
    explicit Parser(std::string desc) : description(std::move(desc)) {}

    void value(std::string const& name, std::string const& help, std::string const& def) {
        options.push_back(Option{name, help, def, false, "", false});
    }
    void flag(std::string const& name, std::string const& help) {
        options.push_back(Option{name, help, "", true, "", false});
    }

    void usage(std::ostream& os) const {
        os << description << "\n\n  OPTIONS:\n\n      -h, --help                        Display this help menu\n";
        for(auto const& o : options) {
            std::string lhs = "      --" + o.name + (o.flag ? "" : "=[" + o.name + "]");
            lhs.resize(std::max<std::size_t>(lhs.size() + 1, 40), ' ');
            os << lhs << o.help;
            if(!o.flag && !o.def.empty())
                os << " Default: " << o.def;
            os << '\n';
        }
    }

    // exits with 0 after --help and with 1 on a malformed command line, as the reference programs
    void parse(int argc, char const* const argv[]) {
        for(int k = 1; k < argc; ++k) {
            std::string a = argv[k];
            if(a == "-h" || a == "--help") {
                usage(std::cout);
                std::exit(0);
            }
            if(a.size() < 3 || a.compare(0, 2, "--") != 0) {
                std::cerr << "Unexpected argument " << a << '\n';
                usage(std::cerr);
                std::exit(1);
            }
            std::string name = a.substr(2), val;
            bool has_val = false;
            auto eq = name.find('=');
            if(eq != std::string::npos) {
                val = name.substr(eq + 1);
                name = name.substr(0, eq);
                has_val = true;
            }
            Option* o = nullptr;
            for(auto& x : options)
                if(x.name == name)
                    o = &x;
            if(!o) {
                std::cerr << "Flag could not be matched: " << name << '\n';
                usage(std::cerr);
                std::exit(1);
            }
            if(o->flag) {
                o->given = true;
                continue;
            }
            if(!has_val) {
                if(k + 1 >= argc) {
                    std::cerr << "Flag '" << name << "' requires an argument but received none\n";
                    usage(std::cerr);
                    std::exit(1);
                }
                val = argv[++k];
            }
            o->value = val;
            o->given = true;
        }
    }

    bool get_flag(std::string const& name) const { return find(name).given; }
    double get_double(std::string const& name) const { return std::stod(text(name)); }
    int get_int(std::string const& name) const { return std::stoi(text(name)); }
    template <typename T> std::vector<T> get_list(std::string const& name) const {
        std::vector<T> out;
        std::stringstream ss(text(name));
        std::string tok;
        while(std::getline(ss, tok, ',')) {
            if(tok.empty())
                continue;
            T x{};
            std::stringstream(tok) >> x;
            out.push_back(x);
        }
        return out;
    }
};

} // namespace progargs
