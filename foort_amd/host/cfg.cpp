// cfg.cpp -- configuration reader, see cfg.hpp.
#include "cfg.hpp"

#include <cctype>
#include <cerrno>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace foort_host {

const Setting *Setting::child(const std::string &name) const
{
    if (type != Group)
        return nullptr;
    for (const auto &c : children)
        if (c.first == name)
            return c.second.get();
    return nullptr;
}

bool Setting::lookup(const std::string &name, double &v) const
{
    const Setting *c = child(name);
    if (!c || c->type != Float)
        return false;
    v = c->f;
    return true;
}

bool Setting::lookup(const std::string &name, bool &v) const
{
    const Setting *c = child(name);
    if (!c || c->type != Bool)
        return false;
    v = c->b;
    return true;
}

bool Setting::lookup(const std::string &name, std::string &v) const
{
    const Setting *c = child(name);
    if (!c || c->type != String)
        return false;
    v = c->s;
    return true;
}

bool Setting::lookup(const std::string &name, int &v) const
{
    const Setting *c = child(name);
    if (!c || c->type != Int)
        return false;
    v = static_cast<int>(c->i);
    return true;
}

bool Setting::lookup_counter(const std::string &name, unsigned long &v) const
{
    const Setting *c = child(name);
    if (!c)
        return false;
    if (c->type == Int64) {
        if (c->i >= 0)
            v = static_cast<unsigned long>(c->i);
        return true;
    }
    if (c->type != Int)
        return false;
    v = c->i < 0 ? 0UL : static_cast<unsigned long>(c->i);
    return true;
}

namespace {

class Parser {
public:
    explicit Parser(const std::string &t) : s(t) {}

    std::unique_ptr<Setting> document()
    {
        auto root = std::make_unique<Setting>();
        root->type = Setting::Group;
        members(*root, '\0');
        skip();
        if (pos < s.size())
            fail("unexpected character '" + std::string(1, s[pos]) + "'");
        return root;
    }

private:
    const std::string &s;
    size_t pos = 0;
    int line = 1;

    [[noreturn]] void fail(const std::string &m) { throw ParseError{line, m}; }

    char peek() const { return pos < s.size() ? s[pos] : '\0'; }

    void adv()
    {
        if (s[pos] == '\n')
            ++line;
        ++pos;
    }

    void skip()
    {
        while (pos < s.size()) {
            char c = s[pos];
            if (std::isspace(static_cast<unsigned char>(c))) {
                adv();
            } else if (c == '#' || (c == '/' && pos + 1 < s.size() && s[pos + 1] == '/')) {
                while (pos < s.size() && s[pos] != '\n')
                    adv();
            } else if (c == '/' && pos + 1 < s.size() && s[pos + 1] == '*') {
                adv();
                adv();
                while (pos < s.size() && !(s[pos] == '*' && pos + 1 < s.size() && s[pos + 1] == '/'))
                    adv();
                if (pos >= s.size())
                    fail("unterminated comment");
                adv();
                adv();
            } else {
                break;
            }
        }
    }

    std::string name()
    {
        skip();
        size_t b = pos;
        if (!(std::isalpha(static_cast<unsigned char>(peek())) || peek() == '*'))
            fail("expected a setting name");
        while (pos < s.size() && (std::isalnum(static_cast<unsigned char>(s[pos])) || s[pos] == '_' ||
                                  s[pos] == '-' || s[pos] == '*'))
            adv();
        return s.substr(b, pos - b);
    }

    // members of a group until `close` ('\0' = end of input)
    void members(Setting &g, char close)
    {
        for (;;) {
            skip();
            if (pos >= s.size()) {
                if (close)
                    fail("missing closing brace");
                return;
            }
            if (peek() == close)
                return;
            std::string n = name();
            skip();
            if (peek() != '=' && peek() != ':')
                fail("expected '=' or ':' after " + n);
            adv();
            auto v = value();
            for (const auto &c : g.children)
                if (c.first == n)
                    fail("duplicate setting " + n);
            g.children.emplace_back(n, std::move(v));
            skip();
            if (peek() == ';' || peek() == ',')
                adv();
        }
    }

    std::unique_ptr<Setting> value()
    {
        skip();
        auto v = std::make_unique<Setting>();
        char c = peek();
        if (c == '{') {
            adv();
            v->type = Setting::Group;
            members(*v, '}');
            adv();
        } else if (c == '(' || c == '[') {
            char close = c == '(' ? ')' : ']';
            adv();
            v->type = Setting::List;
            skip();
            while (peek() != close) {
                v->items.push_back(value());
                skip();
                if (peek() == ',') {
                    adv();
                    skip();
                } else if (peek() != close) {
                    fail("expected ',' in list");
                }
            }
            adv();
        } else if (c == '"') {
            v->type = Setting::String;
            v->s = str();
            // adjacent string literals concatenate
            for (;;) {
                size_t save = pos;
                int sl = line;
                skip();
                if (peek() == '"') {
                    v->s += str();
                } else {
                    pos = save;
                    line = sl;
                    break;
                }
            }
        } else if (std::isalpha(static_cast<unsigned char>(c))) {
            size_t b = pos;
            while (pos < s.size() && std::isalpha(static_cast<unsigned char>(s[pos])))
                adv();
            std::string w = s.substr(b, pos - b);
            for (auto &ch : w)
                ch = static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
            if (w == "true" || w == "false") {
                v->type = Setting::Bool;
                v->b = w == "true";
            } else {
                fail("unexpected word '" + w + "'");
            }
        } else {
            number(*v);
        }
        return v;
    }

    std::string str()
    {
        adv(); // opening quote
        std::string out;
        while (pos < s.size() && s[pos] != '"') {
            char ch = s[pos];
            if (ch == '\\' && pos + 1 < s.size()) {
                adv();
                char e = s[pos];
                out += e == 'n' ? '\n' : e == 't' ? '\t' : e == 'r' ? '\r' : e == 'f' ? '\f' : e;
            } else {
                out += ch;
            }
            adv();
        }
        if (pos >= s.size())
            fail("unterminated string");
        adv();
        return out;
    }

    void number(Setting &v)
    {
        size_t b = pos;
        if (peek() == '+' || peek() == '-')
            adv();
        bool hex = false, flt = false;
        if (peek() == '0' && pos + 1 < s.size() && (s[pos + 1] == 'x' || s[pos + 1] == 'X')) {
            hex = true;
            adv();
            adv();
            while (pos < s.size() && std::isxdigit(static_cast<unsigned char>(s[pos])))
                adv();
        } else {
            while (pos < s.size()) {
                char ch = s[pos];
                if (std::isdigit(static_cast<unsigned char>(ch))) {
                    adv();
                } else if (ch == '.') {
                    flt = true;
                    adv();
                } else if (ch == 'e' || ch == 'E') {
                    flt = true;
                    adv();
                    if (peek() == '+' || peek() == '-')
                        adv();
                } else {
                    break;
                }
            }
        }
        std::string tok = s.substr(b, pos - b);
        if (tok.empty() || tok == "+" || tok == "-")
            fail("expected a value");
        // integer suffix L / LL marks a 64-bit integer
        bool wide = false;
        while (!flt && (peek() == 'L' || peek() == 'l')) {
            wide = true;
            adv();
        }
        errno = 0;
        if (flt) {
            v.type = Setting::Float;
            v.f = std::strtod(tok.c_str(), nullptr);
        } else {
            v.i = std::strtoll(tok.c_str(), nullptr, hex ? 16 : 10);
            if (errno == ERANGE)
                fail("integer out of range");
            wide = wide || v.i < -2147483648LL || v.i > 2147483647LL;
            v.type = wide ? Setting::Int64 : Setting::Int;
        }
    }
};

} // namespace

std::unique_ptr<Setting> parse_config(const std::string &text)
{
    Parser p(text);
    return p.document();
}

std::unique_ptr<Setting> parse_config_file(const std::string &path)
{
    std::ifstream f(path);
    if (!f)
        throw ParseError{0, "cannot read " + path};
    std::stringstream ss;
    ss << f.rdbuf();
    return parse_config(ss.str());
}

} // namespace foort_host
