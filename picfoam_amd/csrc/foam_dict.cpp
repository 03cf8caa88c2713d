// OpenFOAM ASCII dictionary and list reader (see foam_dict.h).
#include "foam_dict.h"

#include <sys/stat.h>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace picmi {
namespace foam {

namespace {

bool is_delim(char c) {
    return c == '(' || c == ')' || c == '{' || c == '}' || c == '[' || c == ']' || c == ';' || c == '"';
}

bool is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\f' || c == '\v'; }

void classify(Token& tk) {
    const std::string& s = tk.text;
    if (s.empty()) return;
    const char c0 = s[0];
    if (!(c0 == '-' || c0 == '+' || c0 == '.' || (c0 >= '0' && c0 <= '9'))) return;
    const char* b = s.c_str();
    char* e = nullptr;
    errno = 0;
    const double v = std::strtod(b, &e);
    if (e != b + s.size() || errno == ERANGE) return;
    tk.kind = Token::Number;
    tk.num = v;
    tk.integral = s.find_first_of(".eE") == std::string::npos;
}

}  // namespace

std::vector<Token> tokenize(const std::string& text, const std::string& where) {
    std::vector<Token> out;
    const size_t n = text.size();
    size_t i = 0;
    int line = 1;
    while (i < n) {
        const char c = text[i];
        if (c == '\n') { ++line; ++i; continue; }
        if (is_space(c)) { ++i; continue; }
        if (c == '/' && i + 1 < n && text[i + 1] == '/') {
            while (i < n && text[i] != '\n') ++i;
            continue;
        }
        if (c == '/' && i + 1 < n && text[i + 1] == '*') {
            i += 2;
            while (i + 1 < n && !(text[i] == '*' && text[i + 1] == '/')) { if (text[i] == '\n') ++line; ++i; }
            if (i + 1 >= n) throw FoamError(where + ": unterminated comment");
            i += 2;
            continue;
        }
        Token tk;
        tk.line = line;
        if (c == '"') {
            size_t j = i + 1;
            std::string s;
            while (j < n && text[j] != '"') {
                if (text[j] == '\\' && j + 1 < n) ++j;
                if (text[j] == '\n') ++line;
                s.push_back(text[j++]);
            }
            if (j >= n) throw FoamError(where + ":" + std::to_string(tk.line) + ": unterminated string");
            tk.kind = Token::String;
            tk.text = s;
            out.push_back(tk);
            i = j + 1;
            continue;
        }
        if (is_delim(c)) {
            tk.kind = Token::Punct;
            tk.text = std::string(1, c);
            out.push_back(tk);
            ++i;
            continue;
        }
        size_t j = i;
        while (j < n && !is_space(text[j]) && !is_delim(text[j]) &&
               !(text[j] == '/' && j + 1 < n && (text[j + 1] == '/' || text[j + 1] == '*')))
            ++j;
        tk.kind = Token::Word;
        tk.text = text.substr(i, j - i);
        classify(tk);
        out.push_back(tk);
        i = j;
    }
    return out;
}

// ---------------------------------------------------------------- Reader

Reader::Reader(const std::vector<Token>& t, std::string where, size_t pos, size_t end)
    : t_(t), where_(std::move(where)), pos_(pos), end_(end > t.size() ? t.size() : end) {}

void Reader::fail(const std::string& what) const {
    std::string at = where_;
    if (pos_ < t_.size()) at += ":" + std::to_string(t_[pos_].line);
    else if (!t_.empty()) at += ":" + std::to_string(t_.back().line);
    throw FoamError(at + ": " + what);
}

const Token& Reader::peek() const {
    if (pos_ >= end_) fail("unexpected end of input");
    return t_[pos_];
}

const Token& Reader::next() {
    const Token& tk = peek();
    ++pos_;
    return tk;
}

void Reader::expect(char c) {
    const Token& tk = next();
    if (!tk.is(c)) { --pos_; fail(std::string("expected '") + c + "', found '" + tk.text + "'"); }
}

bool Reader::accept(char c) {
    if (pos_ < end_ && t_[pos_].is(c)) { ++pos_; return true; }
    return false;
}

double Reader::number() {
    const Token& tk = next();
    if (tk.kind != Token::Number) { --pos_; fail("expected a number, found '" + tk.text + "'"); }
    return tk.num;
}

long long Reader::label() {
    const Token& tk = next();
    if (tk.kind != Token::Number || !tk.integral) { --pos_; fail("expected an integer, found '" + tk.text + "'"); }
    return std::llround(tk.num);
}

std::string Reader::word() {
    const Token& tk = next();
    if (tk.kind != Token::Word && tk.kind != Token::String) { --pos_; fail("expected a word, found '" + tk.text + "'"); }
    return tk.text;
}

std::array<double, 3> Reader::vector3() {
    expect('(');
    std::array<double, 3> v{};
    for (int d = 0; d < 3; ++d) v[d] = number();
    expect(')');
    return v;
}

long long Reader::list_open() {
    long long n = -1;
    if (pos_ < end_ && t_[pos_].kind == Token::Number) n = label();
    expect('(');
    return n;
}

// ---------------------------------------------------------------- Dict

void Dict::fail(const std::string& key, const std::string& what) const {
    throw FoamError(where + ": " + key + ": " + what);
}

const Entry* Dict::find(const std::string& key) const {
    // the last definition wins, as in OpenFOAM's default merge mode
    for (size_t i = entries.size(); i-- > 0;)
        if (entries[i].key == key) return &entries[i];
    return nullptr;
}

const Dict* Dict::sub(const std::string& key) const {
    const Entry* e = find(key);
    return e ? e->dict.get() : nullptr;
}

const Dict& Dict::sub_required(const std::string& key) const {
    const Dict* d = sub(key);
    if (!d) fail(key, "sub-dictionary missing");
    return *d;
}

Reader Dict::value(const std::string& key) const {
    const Entry* e = find(key);
    if (!e) fail(key, "keyword missing");
    if (e->dict) fail(key, "is a dictionary, a value was expected");
    return Reader(e->value, where + ": " + key);
}

double Dict::scalar(const std::string& key) const {
    Reader r = value(key);
    const double v = r.number();
    if (!r.done()) r.fail("a single number was expected");
    return v;
}

double Dict::scalar_or(const std::string& key, double def) const { return has(key) ? scalar(key) : def; }

long long Dict::label(const std::string& key) const {
    Reader r = value(key);
    const long long v = r.label();
    if (!r.done()) r.fail("a single integer was expected");
    return v;
}

std::string Dict::word(const std::string& key) const {
    Reader r = value(key);
    std::string w = r.word();
    if (!r.done()) r.fail("a single word was expected");
    return w;
}

std::string Dict::word_or(const std::string& key, const std::string& def) const { return has(key) ? word(key) : def; }

bool Dict::boolean(const std::string& key) const {
    const std::string w = word(key);
    if (w == "true" || w == "on" || w == "yes" || w == "y" || w == "t") return true;
    if (w == "false" || w == "off" || w == "no" || w == "n" || w == "f" || w == "none") return false;
    fail(key, "'" + w + "' is not a switch (true/false, on/off, yes/no)");
}

bool Dict::boolean_or(const std::string& key, bool def) const { return has(key) ? boolean(key) : def; }

std::array<double, 3> Dict::vector3(const std::string& key) const {
    Reader r = value(key);
    auto v = r.vector3();
    if (!r.done()) r.fail("a single vector was expected");
    return v;
}

std::vector<std::string> Dict::words(const std::string& key) const {
    Reader r = value(key);
    const long long n = r.list_open();
    std::vector<std::string> out;
    while (!r.accept(')')) out.push_back(r.word());
    if (n >= 0 && (long long)out.size() != n) r.fail("list size differs from its count");
    if (!r.done()) r.fail("a single list was expected");
    return out;
}

std::vector<double> Dict::scalars(const std::string& key) const {
    Reader r = value(key);
    const long long n = r.list_open();
    std::vector<double> out;
    while (!r.accept(')')) out.push_back(r.number());
    if (n >= 0 && (long long)out.size() != n) r.fail("list size differs from its count");
    if (!r.done()) r.fail("a single list was expected");
    return out;
}

std::shared_ptr<Dict> parse_entries(const std::vector<Token>& t, size_t& pos, const std::string& where, bool braced) {
    auto d = std::make_shared<Dict>();
    d->where = where;
    const size_t n = t.size();
    while (true) {
        if (pos >= n) {
            if (braced) throw FoamError(where + ": missing '}'");
            return d;
        }
        const Token& k = t[pos];
        if (k.is('}')) {
            if (!braced) throw FoamError(where + ":" + std::to_string(k.line) + ": unmatched '}'");
            ++pos;
            return d;
        }
        if (k.is(';')) { ++pos; continue; }  // stray separators are allowed
        if (k.kind == Token::Punct)
            throw FoamError(where + ":" + std::to_string(k.line) + ": expected a keyword, found '" + k.text + "'");
        if (k.kind == Token::Word && !k.text.empty() && (k.text[0] == '#' || k.text[0] == '$')) {
            if (k.text == "#inputMode") { pos += 2; continue; }
            throw FoamError(where + ":" + std::to_string(k.line) + ": directive or macro '" + k.text +
                            "' is not supported; expand it in the file");
        }
        Entry e;
        e.key = k.text;
        ++pos;
        if (pos < n && t[pos].is('{')) {
            ++pos;
            e.dict = parse_entries(t, pos, where + "/" + e.key, true);
        } else {
            int depth = 0;
            while (true) {
                if (pos >= n) throw FoamError(where + ": " + e.key + ": missing ';'");
                const Token& v = t[pos];
                if (depth == 0 && v.is(';')) { ++pos; break; }
                if (depth == 0 && v.is('}'))
                    throw FoamError(where + ":" + std::to_string(v.line) + ": " + e.key + ": missing ';'");
                if (v.is('(') || v.is('[') || v.is('{')) ++depth;
                if (v.is(')') || v.is(']') || v.is('}')) {
                    if (--depth < 0) throw FoamError(where + ":" + std::to_string(v.line) + ": unbalanced '" + v.text + "'");
                }
                e.value.push_back(v);
                ++pos;
            }
        }
        d->entries.push_back(std::move(e));
    }
}

bool file_exists(const std::string& path) {
    struct stat st;
    return stat(path.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

std::string read_text(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw FoamError(path + ": cannot open");
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

std::shared_ptr<Dict> read_dict_file(const std::string& path) {
    const std::vector<Token> t = tokenize(read_text(path), path);
    size_t pos = 0;
    return parse_entries(t, pos, path, false);
}

ListFile read_list_file(const std::string& path) {
    ListFile lf;
    lf.where = path;
    lf.tokens = tokenize(read_text(path), path);
    size_t pos = 0;
    if (pos + 1 < lf.tokens.size() && lf.tokens[pos].kind == Token::Word && lf.tokens[pos].text == "FoamFile" &&
        lf.tokens[pos + 1].is('{')) {
        pos += 2;
        lf.header = parse_entries(lf.tokens, pos, path + "/FoamFile", true);
        const Dict& h = *lf.header;
        if (h.has("format") && h.word("format") != "ascii")
            throw FoamError(path + ": only ascii files are read (format " + h.word("format") + ")");
    }
    lf.body = pos;
    return lf;
}

std::string header(const std::string& cls, const std::string& location, const std::string& object) {
    std::string s;
    s += "/*--------------------------------*- C++ -*----------------------------------*\\\n";
    s += "  written by libpicmi\n";
    s += "\\*---------------------------------------------------------------------------*/\n";
    s += "FoamFile\n{\n    version     2.0;\n    format      ascii;\n";
    s += "    class       " + cls + ";\n";
    if (!location.empty()) s += "    location    \"" + location + "\";\n";
    s += "    object      " + object + ";\n}\n";
    s += "// * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * //\n\n";
    return s;
}

std::string fmt(double v, int precision) {
    char buf[40];
    std::snprintf(buf, sizeof buf, "%.*g", precision, v);
    return buf;
}

}  // namespace foam
}  // namespace picmi
