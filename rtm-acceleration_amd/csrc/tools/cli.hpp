// cli.hpp -- one declarative option table for the command-line tools.
//
// Each tool declares its options once (name, value kind, destination, help
// line, section); the parser, the "was this set on the command line" query
// that the velocity-header precedence needs, and the usage text all come from
// that one table.  Errors are std::runtime_error, which every tool turns into
// exit status 1.
#pragma once

#include <cstdio>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

namespace rtm {
namespace cli {

class Options {
public:
    explicit Options(std::string synopsis) : synopsis_(std::move(synopsis)) {}

    // Options declared after this call are listed under `heading`.
    Options& section(const std::string& heading) {
        section_ = heading;
        return *this;
    }
    Options& text(const char* name, std::string& dst, const char* help) {
        return add(name, Kind::Text, &dst, help, "TEXT");
    }
    Options& integer(const char* name, int& dst, const char* help) {
        return add(name, Kind::Int, &dst, help, "N");
    }
    Options& real(const char* name, float& dst, const char* help) {
        return add(name, Kind::Real, &dst, help, "X");
    }
    Options& flag(const char* name, bool& dst, const char* help) {
        return add(name, Kind::Flag, &dst, help, "");
    }
    // An option that runs `fn` and ends the program with status 0 (--help,
    // --list-engines); parse() stops at it.
    Options& action(const char* name, std::function<void()> fn, const char* help) {
        add(name, Kind::Action, nullptr, help, "");
        entries_.back().run = std::move(fn);
        return *this;
    }

    // Parses argv[1..argc); returns the arguments that are not options, in
    // order.  Throws on an unknown option or a missing or malformed value.
    std::vector<std::string> parse(int argc, char** argv) {
        std::vector<std::string> positional;
        for (int i = 1; i < argc; ++i) {
            const std::string arg = argv[i];
            if (arg.size() < 2 || arg[0] != '-') {
                positional.push_back(arg);
                continue;
            }
            Entry* e = find(arg);
            if (!e) throw std::runtime_error("unknown option: " + arg);
            e->seen = true;
            if (e->kind == Kind::Action) {
                e->run();
                stopped_ = true;
                return positional;
            }
            if (e->kind == Kind::Flag) {
                *static_cast<bool*>(e->dst) = true;
                continue;
            }
            if (i + 1 >= argc) throw std::runtime_error("missing value after " + arg);
            store(*e, argv[++i]);
        }
        return positional;
    }

    // Whether `name` appeared on the command line.
    bool given(const std::string& name) const {
        for (const auto& e : entries_)
            if (e.name == name) return e.seen;
        throw std::logic_error("no option " + name);
    }
    // Whether an action option ran (the tool should exit 0).
    bool stopped() const { return stopped_; }

    void usage(std::FILE* out) const {
        std::fprintf(out, "%s\n", synopsis_.c_str());
        std::string heading;
        for (const auto& e : entries_) {
            if (e.section != heading) {
                heading = e.section;
                std::fprintf(out, "\n%s:\n", heading.c_str());
            }
            const std::string lhs = e.name + (e.metavar.empty() ? "" : " " + e.metavar);
            std::fprintf(out, "  %-24s %s\n", lhs.c_str(), e.help.c_str());
        }
    }

private:
    enum class Kind { Text, Int, Real, Flag, Action };
    struct Entry {
        std::string name, help, section, metavar;
        Kind kind;
        void* dst;
        std::function<void()> run;
        bool seen = false;
    };

    Options& add(const char* name, Kind kind, void* dst, const char* help, const char* metavar) {
        if (find(name)) throw std::logic_error(std::string("option declared twice: ") + name);
        entries_.push_back({name, help, section_, metavar, kind, dst, {}, false});
        return *this;
    }
    Entry* find(const std::string& name) {
        for (auto& e : entries_)
            if (e.name == name) return &e;
        return nullptr;
    }
    static void store(Entry& e, const std::string& value) {
        std::size_t used = 0;
        try {
            switch (e.kind) {
                case Kind::Text: *static_cast<std::string*>(e.dst) = value; return;
                case Kind::Int: *static_cast<int*>(e.dst) = std::stoi(value, &used); break;
                case Kind::Real: *static_cast<float*>(e.dst) = std::stof(value, &used); break;
                default: return;
            }
        } catch (const std::logic_error&) {
            used = 0;
        }
        if (used != value.size()) throw std::runtime_error("bad value for " + e.name + ": " + value);
    }

    std::string synopsis_, section_ = "Options";
    std::vector<Entry> entries_;
    bool stopped_ = false;
};

}  // namespace cli
}  // namespace rtm
