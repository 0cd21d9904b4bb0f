// Configuration reader/writer: the libconfig subset described in cfg.hpp.
#include "cfg.hpp"

#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace mcm {

namespace {

const char *kind_name(Setting::Kind k) {
    switch (k) {
    case Setting::Group: return "group";
    case Setting::Array: return "array";
    case Setting::List: return "list";
    case Setting::Int: return "integer";
    case Setting::Float: return "float";
    case Setting::Bool: return "boolean";
    case Setting::String: return "string";
    }
    return "?";
}

class Parser {
  public:
    Parser(const std::string &text, const std::string &origin) : s_(text), origin_(origin) {}

    Setting parse_file() {
        Setting root;
        root.kind = Setting::Group;
        root.path = "";
        skip();
        while (pos_ < s_.size()) {
            parse_setting(root);
            skip();
        }
        return root;
    }

  private:
    const std::string &s_;
    std::string origin_;
    size_t pos_ = 0;
    int line_ = 1;

    [[noreturn]] void fail(const std::string &what) const {
        throw CfgError("parse error at " + origin_ + ":" + std::to_string(line_) + " - " + what);
    }

    char peek() const { return pos_ < s_.size() ? s_[pos_] : '\0'; }
    char get() {
        char c = s_[pos_++];
        if (c == '\n') ++line_;
        return c;
    }

    void skip() {  // white space and comments
        while (pos_ < s_.size()) {
            char c = s_[pos_];
            if (std::isspace((unsigned char)c)) {
                get();
            } else if (c == '#' || (c == '/' && pos_ + 1 < s_.size() && s_[pos_ + 1] == '/')) {
                while (pos_ < s_.size() && s_[pos_] != '\n') get();
            } else if (c == '/' && pos_ + 1 < s_.size() && s_[pos_ + 1] == '*') {
                get();
                get();
                while (pos_ + 1 < s_.size() && !(s_[pos_] == '*' && s_[pos_ + 1] == '/')) get();
                if (pos_ + 1 >= s_.size()) fail("unterminated comment");
                get();
                get();
            } else {
                break;
            }
        }
    }

    static bool name_start(char c) { return std::isalpha((unsigned char)c) || c == '*'; }
    static bool name_char(char c) { return std::isalnum((unsigned char)c) || c == '*' || c == '_' || c == '-'; }

    void parse_setting(Setting &group) {
        if (!name_start(peek())) fail(std::string("expected a setting name, found '") + peek() + "'");
        std::string name;
        while (pos_ < s_.size() && name_char(peek())) name += get();
        skip();
        if (peek() != '=' && peek() != ':') fail("expected '=' or ':' after '" + name + "'");
        get();
        skip();
        for (const Setting &m : group.items)
            if (m.name == name) fail("duplicate setting '" + name + "'");
        Setting v = parse_value(group.path.empty() ? name : group.path + "." + name);
        v.name = name;
        group.items.push_back(std::move(v));
        skip();
        if (peek() == ';' || peek() == ',') get();
    }

    Setting parse_value(const std::string &path) {
        Setting v;
        v.path = path;
        char c = peek();
        if (c == '{') {
            get();
            v.kind = Setting::Group;
            skip();
            while (peek() != '}') {
                if (pos_ >= s_.size()) fail("unterminated group '" + path + "'");
                parse_setting(v);
                skip();
            }
            get();
        } else if (c == '[' || c == '(') {
            const char close = c == '[' ? ']' : ')';
            get();
            v.kind = c == '[' ? Setting::Array : Setting::List;
            skip();
            while (peek() != close) {
                if (pos_ >= s_.size()) fail(std::string("unterminated ") + kind_name(v.kind) + " '" + path + "'");
                Setting e = parse_value(path + ".[" + std::to_string(v.items.size()) + "]");  // libconfig's path form
                if (v.kind == Setting::Array) {
                    if (e.is_aggregate()) fail("an array holds scalars only ('" + path + "')");
                    if (!v.items.empty() && v.items[0].kind != e.kind &&
                        !(v.items[0].is_number() && e.is_number()))
                        fail("mixed types in array '" + path + "'");
                }
                v.items.push_back(std::move(e));
                skip();
                if (peek() == ',') {
                    get();
                    skip();
                } else if (peek() != close) {
                    fail(std::string("expected ',' or '") + close + "' in '" + path + "'");
                }
            }
            get();
            if (v.kind == Setting::Array) {  // an array of mixed int/float holds floats
                bool any_float = false;
                for (const Setting &e : v.items) any_float |= e.kind == Setting::Float;
                if (any_float)
                    for (Setting &e : v.items)
                        if (e.kind == Setting::Int) {
                            e.kind = Setting::Float;
                            e.fval = (double)e.ival;
                        }
            }
        } else if (c == '"') {
            v.kind = Setting::String;
            while (peek() == '"') {
                get();
                while (peek() != '"') {
                    if (pos_ >= s_.size() || peek() == '\n') fail("unterminated string");
                    char ch = get();
                    if (ch == '\\') {
                        char e = get();
                        switch (e) {
                        case 'n': ch = '\n'; break;
                        case 't': ch = '\t'; break;
                        case 'r': ch = '\r'; break;
                        case 'f': ch = '\f'; break;
                        case '\\': ch = '\\'; break;
                        case '"': ch = '"'; break;
                        default: fail(std::string("unknown escape '\\") + e + "'");
                        }
                    }
                    v.sval += ch;
                }
                get();
                skip();
            }
        } else {
            size_t start = pos_;
            while (pos_ < s_.size() && (std::isalnum((unsigned char)peek()) || peek() == '.' || peek() == '+' ||
                                        peek() == '-'))
                get();
            std::string tok = s_.substr(start, pos_ - start);
            if (tok.empty()) fail(std::string("expected a value for '") + path + "'");
            std::string low;
            for (char ch : tok) low += (char)std::tolower((unsigned char)ch);
            if (low == "true" || low == "false") {
                v.kind = Setting::Bool;
                v.bval = low == "true";
                return v;
            }
            bool is_hex = low.size() > 2 && (low.compare(0, 2, "0x") == 0);
            bool is_float = !is_hex && (low.find_first_of(".e") != std::string::npos || low == "nan" ||
                                        low == "inf" || low == "+inf" || low == "-inf");
            errno = 0;
            char *end = nullptr;
            if (is_float) {
                v.kind = Setting::Float;
                v.fval = std::strtod(tok.c_str(), &end);
            } else {
                std::string t = tok;
                if (!t.empty() && (t.back() == 'L' || t.back() == 'l')) t.pop_back();
                if (!t.empty() && (t.back() == 'L' || t.back() == 'l')) t.pop_back();
                v.kind = Setting::Int;
                v.ival = std::strtoll(t.c_str(), &end, 0);
                if (*end == '\0' && errno == 0) return v;
                fail("bad number '" + tok + "' for '" + path + "'");
            }
            if (end == nullptr || *end != '\0' || errno == ERANGE) fail("bad number '" + tok + "' for '" + path + "'");
        }
        return v;
    }
};

void format_value(const Setting &v, std::ostringstream &o, int indent);

void format_scalar(const Setting &v, std::ostringstream &o) {
    char buf[64];
    switch (v.kind) {
    case Setting::Int: o << v.ival; break;
    case Setting::Float:
        std::snprintf(buf, sizeof buf, "%.17g", v.fval);
        o << buf;
        if (std::isfinite(v.fval) && !std::strpbrk(buf, ".e")) o << ".0";
        break;
    case Setting::Bool: o << (v.bval ? "true" : "false"); break;
    case Setting::String:
        o << '"';
        for (char c : v.sval) {
            if (c == '"' || c == '\\') o << '\\' << c;
            else if (c == '\n') o << "\\n";
            else if (c == '\t') o << "\\t";
            else o << c;
        }
        o << '"';
        break;
    default: break;
    }
}

void format_group_members(const Setting &g, std::ostringstream &o, int indent) {
    for (const Setting &m : g.items) {
        o << std::string(indent, ' ') << m.name << " = ";
        format_value(m, o, indent);
        o << ";\n";
    }
}

void format_value(const Setting &v, std::ostringstream &o, int indent) {
    if (v.kind == Setting::Group) {
        o << "{\n";
        format_group_members(v, o, indent + 2);
        o << std::string(indent, ' ') << "}";
    } else if (v.kind == Setting::Array) {
        o << "[ ";
        for (size_t i = 0; i < v.items.size(); ++i) {
            if (i) o << ", ";
            format_scalar(v.items[i], o);
        }
        o << " ]";
    } else if (v.kind == Setting::List) {
        o << "( ";
        for (size_t i = 0; i < v.items.size(); ++i) {
            if (i) o << ", ";
            format_value(v.items[i], o, indent + 2);
        }
        o << " )";
    } else {
        format_scalar(v, o);
    }
}

}  // namespace

bool Setting::exists(const std::string &key) const {
    if (kind != Group) return false;
    for (const Setting &m : items)
        if (m.name == key) return true;
    return false;
}

Setting *Setting::find(const std::string &key) {
    if (kind != Group) return nullptr;
    for (Setting &m : items)
        if (m.name == key) return &m;
    return nullptr;
}

const Setting &Setting::operator[](const std::string &key) const {
    if (kind == Group)
        for (const Setting &m : items)
            if (m.name == key) return m;
    throw CfgError("No " + (path.empty() ? key : path + "." + key) + " setting in configuration file.");
}

const Setting &Setting::operator[](int i) const {
    if (!is_aggregate() || i < 0 || i >= (int)items.size())
        throw CfgError("No " + path + ".[" + std::to_string(i) + "] setting in configuration file.");
    return items[i];
}

double Setting::as_double() const {
    if (kind == Float) return fval;
    if (kind == Int) return (double)ival;
    throw CfgError("setting " + path + " is a " + kind_name(kind) + ", a number was expected");
}

long long Setting::as_int() const {
    if (kind == Int) return ival;
    throw CfgError("setting " + path + " is a " + kind_name(kind) + ", an integer was expected");
}

bool Setting::as_bool() const {
    if (kind == Bool) return bval;
    throw CfgError("setting " + path + " is a " + kind_name(kind) + ", a boolean was expected");
}

const std::string &Setting::as_string() const {
    if (kind == String) return sval;
    throw CfgError("setting " + path + " is a " + kind_name(kind) + ", a string was expected");
}

bool Setting::lookup(const std::string &key, double &out) const {
    if (!exists(key)) return false;
    const Setting &m = (*this)[key];
    if (!m.is_number()) return false;
    out = m.as_double();
    return true;
}

bool Setting::lookup(const std::string &key, std::string &out) const {
    if (!exists(key)) return false;
    const Setting &m = (*this)[key];
    if (m.kind != String) return false;
    out = m.sval;
    return true;
}

void Setting::set_double(const std::string &key, double v) {
    Setting *m = find(key);
    if (!m) throw CfgError("No " + (path.empty() ? key : path + "." + key) + " setting in configuration file.");
    m->kind = Float;
    m->fval = v;
}

Setting cfg_parse(const std::string &text, const std::string &origin) {
    Parser p(text, origin);
    return p.parse_file();
}

Setting cfg_read_file(const std::string &path) {
    std::ifstream in(path);
    if (!in.good()) throw CfgError("I/O error while reading file " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    return cfg_parse(ss.str(), path);
}

std::string cfg_format(const Setting &root) {
    std::ostringstream o;
    format_group_members(root, o, 0);
    return o.str();
}

void cfg_write_file(const Setting &root, const std::string &path) {
    std::string tmp = path + ".tmp";
    {
        std::ofstream out(tmp);
        if (!out.good()) throw CfgError("cannot write " + tmp);
        out << cfg_format(root);
        if (!out.good()) throw CfgError("cannot write " + tmp);
    }
    if (std::rename(tmp.c_str(), path.c_str()) != 0) throw CfgError("cannot replace " + path);
}

}  // namespace mcm
