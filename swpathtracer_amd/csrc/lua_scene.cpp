// lua_scene.cpp - scene scripts of the reference without liblua.
//
// Three parts:
//  1. an interpreter for the Lua 5.3 subset that scene scripts and the
//     reference's lua_core/*.lua use: locals, globals, tables (constructors,
//     fields, indexing, #), arithmetic, comparison, logic, string
//     concatenation, if / while / repeat / numeric and generic for, function
//     definitions and closures, return / break, calls with `f{...}` and
//     `f"..."` sugar, `require` over package.path, and the library functions
//     those scripts touch (print, type, pairs, ipairs, tostring, tonumber,
//     error, assert, os.exit, math.*, string.*, table.*);
//  2. the scene functions of lua_core/*.lua as built-ins, used when
//     `require` finds no lua_core file on package.path: the same defaults,
//     the same validation (reported with the reference's messages) and the
//     same registry layout under _G.__RENDERER_ELEMENTS__
//     (lua_core/registervars.lua:1-16);
//  3. the extraction lua_bind/lua_bind.cpp performs (getNumberField etc.) and
//     the translation SwpathtracerApp does from it (core/swpathtracer_app.cpp:
//     cameras, BSDF samplers, layered BSDFs, materials, primitives, meshes
//     through an OBJ reader with Scene::loadMesh's order, integrator).
//
// Errors: the reference prints "Lua ERROR: ..." and calls os.exit() from the
// validation functions, and throws std::runtime_error from lua_bind; here both
// become a std::runtime_error whose text is the reference's message.
#include "lua_scene.h"

#include <array>
#include <cerrno>
#include <ctime>
#include <filesystem>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <limits>
#include <map>
#include <sstream>
#include <stdexcept>
#include <unordered_map>
#include <unordered_set>

namespace spt_lua {

// ===========================================================================
// glibc random(): r[0] = seed, r[i] = 16807 r[i-1] mod (2^31 - 1) for
// i < 31 (Schrage's method, as srandom computes it), r[i] = r[i-31] for
// 31 <= i < 34, then r[i] = r[i-31] + r[i-3] mod 2^32; output k is
// r[k + 344] >> 1 (srandom discards the first 310 outputs).
// ===========================================================================

void GlibcRandom::seed_with(uint32_t seed) {
    int64_t word = (int32_t)(seed == 0 ? 1u : seed);
    int32_t r[344];
    r[0] = (int32_t)word;
    for (int i = 1; i < 31; ++i) {
        const int64_t hi = word / 127773, lo = word % 127773;
        word = 16807 * lo - 2836 * hi;
        if (word < 0) word += 2147483647;
        r[i] = (int32_t)word;
    }
    for (int i = 31; i < 34; ++i) r[i] = r[i - 31];
    for (int i = 34; i < 344; ++i) r[i] = (int32_t)((uint32_t)r[i - 31] + (uint32_t)r[i - 3]);
    for (int i = 310; i < 344; ++i) r_[i % 34] = r[i];  // r[j] lives at ring slot j mod 34
    k_ = 344;
}

int32_t GlibcRandom::next() {
    const int32_t v = (int32_t)((uint32_t)r_[(k_ - 31) % 34] + (uint32_t)r_[(k_ - 3) % 34]);
    r_[k_ % 34] = v;
    ++k_;
    if (k_ >= 344 + 34 * 1000000) k_ -= 34 * 1000000;  // keep the index small; ring positions unchanged
    return (int32_t)((uint32_t)v >> 1);
}

namespace {

// ===========================================================================
// Values
// ===========================================================================

struct Table;
struct Function;
struct Interp;

struct Value {
    enum Kind : uint8_t { NIL, BOOL, NUM, STR, TABLE, FUNC };
    Kind k = NIL;
    bool b = false;
    bool isint = false;  // Lua 5.3's integer subtype (exact up to 2^53 here)
    double n = 0.0;
    std::shared_ptr<const std::string> s;
    std::shared_ptr<Table> t;
    std::shared_ptr<Function> f;

    static Value nil() { return Value(); }
    static Value boolean(bool v) { Value x; x.k = BOOL; x.b = v; return x; }
    static Value number(double v) { Value x; x.k = NUM; x.n = v; return x; }
    static Value integer(double v) { Value x; x.k = NUM; x.n = v; x.isint = true; return x; }
    static Value string(std::string v) {
        Value x; x.k = STR; x.s = std::make_shared<const std::string>(std::move(v)); return x;
    }
    bool truthy() const { return !(k == NIL || (k == BOOL && !b)); }
    bool is_nil() const { return k == NIL; }
};

using Values = std::vector<Value>;

const char* type_name(const Value& v) {
    switch (v.k) {
        case Value::NIL: return "nil";
        case Value::BOOL: return "boolean";
        case Value::NUM: return "number";
        case Value::STR: return "string";
        case Value::TABLE: return "table";
        default: return "function";
    }
}

bool raw_equal(const Value& a, const Value& b) {
    if (a.k != b.k) return false;
    switch (a.k) {
        case Value::NIL: return true;
        case Value::BOOL: return a.b == b.b;
        case Value::NUM: return a.n == b.n;
        case Value::STR: return a.s == b.s || *a.s == *b.s;
        case Value::TABLE: return a.t == b.t;
        default: return a.f == b.f;
    }
}

// Table key: numbers with an integral value are integer keys (Lua 5.3
// normalises 2.0 to 2), strings by content, the rest by identity.
struct Key {
    uint8_t kind = 0;  // 1 bool, 2 number, 3 string, 4 table, 5 function
    double n = 0.0;
    std::string s;
    const void* p = nullptr;
    bool operator==(const Key& o) const {
        return kind == o.kind && n == o.n && s == o.s && p == o.p;
    }
};
struct KeyHash {
    size_t operator()(const Key& k) const {
        size_t h = std::hash<uint8_t>()(k.kind);
        h ^= std::hash<double>()(k.n) + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
        h ^= std::hash<std::string>()(k.s) + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
        h ^= std::hash<const void*>()(k.p) + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
        return h;
    }
};

struct Table {
    std::vector<Value> arr;  // arr[i] is key i + 1
    std::unordered_map<Key, size_t, KeyHash> index;
    std::vector<std::pair<Value, Value>> entries;  // hash part in insertion order (erased: key nil)

    static bool array_index(const Value& key, size_t& i) {
        if (key.k != Value::NUM) return false;
        const double d = key.n;
        if (!(d >= 1.0) || d > 9.0e15 || std::floor(d) != d) return false;
        i = (size_t)d - 1;
        return true;
    }
    static Key make_key(const Value& v) {
        Key k;
        switch (v.k) {
            case Value::BOOL: k.kind = 1; k.n = v.b ? 1.0 : 0.0; break;
            case Value::NUM: k.kind = 2; k.n = v.n == 0.0 ? 0.0 : v.n; break;
            case Value::STR: k.kind = 3; k.s = *v.s; break;
            case Value::TABLE: k.kind = 4; k.p = v.t.get(); break;
            case Value::FUNC: k.kind = 5; k.p = v.f.get(); break;
            default: break;
        }
        return k;
    }
    Value get(const Value& key) const {
        size_t i;
        if (array_index(key, i) && i < arr.size()) return arr[i];
        if (key.k == Value::NIL) return Value();
        auto it = index.find(make_key(key));
        return it == index.end() ? Value() : entries[it->second].second;
    }
    Value get(const std::string& name) const { return get(Value::string(name)); }
    void set(const Value& key, const Value& val) {
        size_t i;
        if (array_index(key, i)) {
            if (i < arr.size()) {
                arr[i] = val;
                if (val.is_nil() && i + 1 == arr.size())
                    while (!arr.empty() && arr.back().is_nil()) arr.pop_back();
                return;
            }
            if (i == arr.size() && !val.is_nil()) {
                erase_hash(key);
                arr.push_back(val);
                // migrate the keys that now continue the array part
                for (;;) {
                    Value next = Value::integer((double)arr.size() + 1.0);
                    auto it = index.find(make_key(next));
                    if (it == index.end() || entries[it->second].second.is_nil()) break;
                    arr.push_back(entries[it->second].second);
                    erase_hash(next);
                }
                return;
            }
        }
        const Key k = make_key(key);
        auto it = index.find(k);
        if (val.is_nil()) {
            if (it != index.end()) erase_hash(key);
            return;
        }
        if (it != index.end()) {
            entries[it->second].second = val;
        } else {
            index.emplace(k, entries.size());
            entries.emplace_back(key, val);
        }
    }
    void set(const std::string& name, const Value& val) { set(Value::string(name), val); }
    // A removed hash entry keeps its key and position with a nil value, so
    // a traversal (next) that clears fields as it goes can continue from it.
    void erase_hash(const Value& key) {
        auto it = index.find(make_key(key));
        if (it == index.end()) return;
        entries[it->second].second = Value();
    }
    bool has_hash(const Value& key) const {
        auto it = index.find(make_key(key));
        return it != index.end() && !entries[it->second].second.is_nil();
    }
    // The border # returns: the array part's length, extended through the
    // hash part while consecutive integer keys are present.
    size_t length() const {
        size_t n = arr.size();
        while (n > 0 && arr[n - 1].is_nil()) --n;
        if (n == arr.size())
            while (has_hash(Value::integer((double)n + 1.0))) ++n;
        return n;
    }
    // Iteration order of next/pairs: the array part, then the hash part in
    // insertion order.
    bool next(const Value& key, Value& k_out, Value& v_out) const {
        size_t pos = 0;  // position in the combined sequence to start from
        if (!key.is_nil()) {
            size_t i;
            auto it = index.find(make_key(key));
            if (array_index(key, i) && i < arr.size()) {
                pos = i + 1;
            } else if (it != index.end()) {
                pos = arr.size() + it->second + 1;
            } else if (array_index(key, i)) {
                pos = arr.size();  // the array part shrank under the traversal
            } else {
                throw std::runtime_error("invalid key to 'next'");
            }
        }
        for (; pos < arr.size(); ++pos)
            if (!arr[pos].is_nil()) {
                k_out = Value::integer((double)pos + 1.0);
                v_out = arr[pos];
                return true;
            }
        for (size_t e = pos - arr.size(); e < entries.size(); ++e)
            if (!entries[e].second.is_nil()) {
                k_out = entries[e].first;
                v_out = entries[e].second;
                return true;
            }
        return false;
    }
};

// ===========================================================================
// Lexer
// ===========================================================================

enum Tok {
    T_EOF, T_NAME, T_NUMBER, T_STRING,
    // keywords
    T_AND, T_BREAK, T_DO, T_ELSE, T_ELSEIF, T_END, T_FALSE, T_FOR, T_FUNCTION, T_GOTO, T_IF, T_IN,
    T_LOCAL, T_NIL, T_NOT, T_OR, T_REPEAT, T_RETURN, T_THEN, T_TRUE, T_UNTIL, T_WHILE,
    // symbols
    T_PLUS, T_MINUS, T_STAR, T_SLASH, T_DSLASH, T_PERCENT, T_CARET, T_HASH, T_EQ, T_NE, T_LE, T_GE, T_LT,
    T_GT, T_ASSIGN, T_LPAREN, T_RPAREN, T_LBRACE, T_RBRACE, T_LBRACKET, T_RBRACKET, T_SEMI, T_COLON,
    T_COMMA, T_DOT, T_CONCAT, T_ELLIPSIS
};

struct Token {
    Tok t = T_EOF;
    std::string s;
    double n = 0.0;
    bool isint = false;
    int line = 1;
};

class Lexer {
   public:
    Lexer(const std::string& src, const std::string& chunk) : src_(src), chunk_(chunk) {}
    std::vector<Token> run() {
        std::vector<Token> out;
        for (;;) {
            Token t = next();
            out.push_back(t);
            if (t.t == T_EOF) break;
        }
        return out;
    }

   private:
    [[noreturn]] void fail(const std::string& msg) {
        throw std::runtime_error(chunk_ + ":" + std::to_string(line_) + ": " + msg);
    }
    char peek(size_t o = 0) const { return p_ + o < src_.size() ? src_[p_ + o] : '\0'; }
    // [[ ... ]] / [==[ ... ]==]: returns false when the text at p_ is not one
    bool long_bracket(std::string* out) {
        size_t q = p_ + 1;
        int level = 0;
        while (q < src_.size() && src_[q] == '=') ++level, ++q;
        if (q >= src_.size() || src_[q] != '[') return false;
        q++;
        if (q < src_.size() && src_[q] == '\r') ++q;
        if (q < src_.size() && src_[q] == '\n') ++q, ++line_;
        const std::string close = "]" + std::string(level, '=') + "]";
        const size_t e = src_.find(close, q);
        if (e == std::string::npos) fail("unfinished long string/comment");
        for (size_t i = q; i < e; ++i)
            if (src_[i] == '\n') ++line_;
        if (out) *out = src_.substr(q, e - q);
        p_ = e + close.size();
        return true;
    }
    Token next() {
        for (;;) {
            const char c = peek();
            if (c == '\n') { ++line_; ++p_; continue; }
            if (c == ' ' || c == '\t' || c == '\r' || c == '\f' || c == '\v') { ++p_; continue; }
            if (c == '-' && peek(1) == '-') {
                p_ += 2;
                if (peek() == '[' && long_bracket(nullptr)) continue;
                while (p_ < src_.size() && src_[p_] != '\n') ++p_;
                continue;
            }
            break;
        }
        Token t;
        t.line = line_;
        if (p_ >= src_.size()) return t;
        const char c = peek();
        if (std::isalpha((unsigned char)c) || c == '_') {
            size_t q = p_;
            while (q < src_.size() && (std::isalnum((unsigned char)src_[q]) || src_[q] == '_')) ++q;
            t.s = src_.substr(p_, q - p_);
            p_ = q;
            static const std::unordered_map<std::string, Tok> kw = {
                {"and", T_AND}, {"break", T_BREAK}, {"do", T_DO}, {"else", T_ELSE}, {"elseif", T_ELSEIF},
                {"end", T_END}, {"false", T_FALSE}, {"for", T_FOR}, {"function", T_FUNCTION}, {"goto", T_GOTO},
                {"if", T_IF}, {"in", T_IN}, {"local", T_LOCAL}, {"nil", T_NIL}, {"not", T_NOT}, {"or", T_OR},
                {"repeat", T_REPEAT}, {"return", T_RETURN}, {"then", T_THEN}, {"true", T_TRUE},
                {"until", T_UNTIL}, {"while", T_WHILE}};
            auto it = kw.find(t.s);
            t.t = it == kw.end() ? T_NAME : it->second;
            return t;
        }
        if (std::isdigit((unsigned char)c) || (c == '.' && std::isdigit((unsigned char)peek(1)))) {
            size_t q = p_;
            if (c == '0' && (peek(1) == 'x' || peek(1) == 'X')) {
                q += 2;
                while (q < src_.size() && std::isxdigit((unsigned char)src_[q])) ++q;
                t.n = (double)(int64_t)std::strtoull(src_.substr(p_ + 2, q - p_ - 2).c_str(), nullptr, 16);
                t.isint = true;
            } else {
                while (q < src_.size() && (std::isdigit((unsigned char)src_[q]) || src_[q] == '.')) ++q;
                if (q < src_.size() && (src_[q] == 'e' || src_[q] == 'E')) {
                    ++q;
                    if (q < src_.size() && (src_[q] == '+' || src_[q] == '-')) ++q;
                    while (q < src_.size() && std::isdigit((unsigned char)src_[q])) ++q;
                }
                const std::string num = src_.substr(p_, q - p_);
                char* end = nullptr;
                t.n = std::strtod(num.c_str(), &end);
                if (!end || *end) fail("malformed number near '" + num + "'");
                t.isint = num.find_first_of(".eE") == std::string::npos && std::fabs(t.n) < 9.2e18;
            }
            p_ = q;
            t.t = T_NUMBER;
            return t;
        }
        if (c == '"' || c == '\'') {
            ++p_;
            std::string s;
            for (;;) {
                if (p_ >= src_.size() || src_[p_] == '\n') fail("unfinished string");
                const char d = src_[p_++];
                if (d == c) break;
                if (d != '\\') { s += d; continue; }
                const char e = peek();
                ++p_;
                switch (e) {
                    case 'n': s += '\n'; break;
                    case 't': s += '\t'; break;
                    case 'r': s += '\r'; break;
                    case 'a': s += '\a'; break;
                    case 'b': s += '\b'; break;
                    case 'f': s += '\f'; break;
                    case 'v': s += '\v'; break;
                    case '\\': s += '\\'; break;
                    case '"': s += '"'; break;
                    case '\'': s += '\''; break;
                    case '\n': s += '\n'; ++line_; break;
                    default:
                        if (std::isdigit((unsigned char)e)) {
                            int v = e - '0', k = 1;
                            while (k < 3 && std::isdigit((unsigned char)peek())) v = v * 10 + (src_[p_++] - '0'), ++k;
                            s += (char)v;
                        } else {
                            fail("invalid escape sequence");
                        }
                }
            }
            t.t = T_STRING;
            t.s = s;
            return t;
        }
        if (c == '[' && (peek(1) == '[' || peek(1) == '=')) {
            std::string s;
            if (long_bracket(&s)) {
                t.t = T_STRING;
                t.s = s;
                return t;
            }
        }
        struct Sym { const char* s; Tok t; };
        static const Sym syms[] = {
            {"...", T_ELLIPSIS}, {"..", T_CONCAT}, {"==", T_EQ}, {"~=", T_NE}, {"<=", T_LE}, {">=", T_GE},
            {"//", T_DSLASH}, {"+", T_PLUS}, {"-", T_MINUS}, {"*", T_STAR}, {"/", T_SLASH}, {"%", T_PERCENT},
            {"^", T_CARET}, {"#", T_HASH}, {"<", T_LT}, {">", T_GT}, {"=", T_ASSIGN}, {"(", T_LPAREN},
            {")", T_RPAREN}, {"{", T_LBRACE}, {"}", T_RBRACE}, {"[", T_LBRACKET}, {"]", T_RBRACKET},
            {";", T_SEMI}, {":", T_COLON}, {",", T_COMMA}, {".", T_DOT}};
        for (const Sym& sy : syms) {
            const size_t len = std::strlen(sy.s);
            if (src_.compare(p_, len, sy.s) == 0) {
                p_ += len;
                t.t = sy.t;
                t.s = sy.s;
                return t;
            }
        }
        fail(std::string("unexpected symbol near '") + c + "'");
    }

    const std::string& src_;
    std::string chunk_;
    size_t p_ = 0;
    int line_ = 1;
};

// ===========================================================================
// Syntax tree
// ===========================================================================

struct Expr;
struct Stat;
using ExprP = std::unique_ptr<Expr>;
using StatP = std::unique_ptr<Stat>;
using Block = std::vector<StatP>;

struct FuncBody {
    std::vector<std::string> params;
    bool vararg = false;
    Block body;
    std::string name;
    int line = 0;
};

struct TableField {
    int kind = 0;  // 0 positional, 1 name = e, 2 [k] = e
    std::string name;
    ExprP key, val;
};

struct Expr {
    enum Kind { NIL, TRUE_, FALSE_, NUM, STR, VARARG, FUNC, TABLE, BIN, UN, NAME, INDEX, CALL, METHOD, PAREN, AND, OR };
    Kind k;
    int line = 0;
    double n = 0.0;
    bool isint = false;
    std::string s;  // string literal, name, method name
    Tok op = T_EOF;
    ExprP a, b;
    std::vector<ExprP> args;
    std::vector<TableField> fields;
    std::shared_ptr<FuncBody> fn;
    explicit Expr(Kind kind) : k(kind) {}
    bool multi() const { return k == CALL || k == METHOD || k == VARARG; }
};

struct Stat {
    enum Kind { LOCAL, ASSIGN, CALL, DO, WHILE, REPEAT, IF, NUMFOR, GENFOR, FUNC, LOCALFUNC, RETURN, BREAK };
    Kind k;
    int line = 0;
    std::vector<std::string> names;
    std::vector<ExprP> targets, exprs;
    ExprP cond;
    Block body, else_body;
    std::vector<ExprP> conds;
    std::vector<Block> blocks;
    std::shared_ptr<FuncBody> fn;
    explicit Stat(Kind kind) : k(kind) {}
};

// ===========================================================================
// Parser (recursive descent, Lua 5.3 precedence)
// ===========================================================================

class Parser {
   public:
    Parser(std::vector<Token> toks, std::string chunk) : t_(std::move(toks)), chunk_(std::move(chunk)) {}
    std::shared_ptr<FuncBody> chunk() {
        auto f = std::make_shared<FuncBody>();
        f->vararg = true;
        f->name = "main chunk";
        f->body = block();
        if (cur().t != T_EOF) fail("'<eof>' expected");
        return f;
    }

   private:
    const Token& cur() const { return t_[i_]; }
    const Token& look(size_t o) const { return t_[std::min(i_ + o, t_.size() - 1)]; }
    bool accept(Tok t) {
        if (cur().t != t) return false;
        ++i_;
        return true;
    }
    [[noreturn]] void fail(const std::string& msg) const {
        throw std::runtime_error(chunk_ + ":" + std::to_string(cur().line) + ": " + msg + " near '" +
                                 (cur().t == T_EOF ? std::string("<eof>") : cur().s.empty() ? std::string("?") : cur().s) +
                                 "'");
    }
    void expect(Tok t, const char* what) {
        if (!accept(t)) fail(std::string("'") + what + "' expected");
    }
    std::string name() {
        if (cur().t != T_NAME) fail("<name> expected");
        return t_[i_++].s;
    }
    static bool block_end(Tok t) { return t == T_EOF || t == T_END || t == T_ELSE || t == T_ELSEIF || t == T_UNTIL; }

    Block block() {
        Block b;
        while (!block_end(cur().t)) {
            if (cur().t == T_RETURN) {
                b.push_back(ret());
                break;
            }
            StatP s = statement();
            if (s) b.push_back(std::move(s));
        }
        return b;
    }
    StatP ret() {
        auto s = std::make_unique<Stat>(Stat::RETURN);
        s->line = cur().line;
        ++i_;
        if (!block_end(cur().t) && cur().t != T_SEMI) s->exprs = exprlist();
        accept(T_SEMI);
        return s;
    }
    std::shared_ptr<FuncBody> funcbody(const std::string& fname, bool method) {
        auto f = std::make_shared<FuncBody>();
        f->name = fname;
        f->line = cur().line;
        if (method) f->params.push_back("self");
        expect(T_LPAREN, "(");
        if (cur().t != T_RPAREN) {
            do {
                if (accept(T_ELLIPSIS)) {
                    f->vararg = true;
                    break;
                }
                f->params.push_back(name());
            } while (accept(T_COMMA));
        }
        expect(T_RPAREN, ")");
        f->body = block();
        expect(T_END, "end");
        return f;
    }
    StatP statement() {
        const int line = cur().line;
        switch (cur().t) {
            case T_SEMI: ++i_; return nullptr;
            case T_BREAK: { ++i_; auto s = std::make_unique<Stat>(Stat::BREAK); s->line = line; return s; }
            case T_DO: {
                ++i_;
                auto s = std::make_unique<Stat>(Stat::DO);
                s->line = line;
                s->body = block();
                expect(T_END, "end");
                return s;
            }
            case T_WHILE: {
                ++i_;
                auto s = std::make_unique<Stat>(Stat::WHILE);
                s->line = line;
                s->cond = expr();
                expect(T_DO, "do");
                s->body = block();
                expect(T_END, "end");
                return s;
            }
            case T_REPEAT: {
                ++i_;
                auto s = std::make_unique<Stat>(Stat::REPEAT);
                s->line = line;
                s->body = block();
                expect(T_UNTIL, "until");
                s->cond = expr();
                return s;
            }
            case T_IF: {
                ++i_;
                auto s = std::make_unique<Stat>(Stat::IF);
                s->line = line;
                s->conds.push_back(expr());
                expect(T_THEN, "then");
                s->blocks.push_back(block());
                for (;;) {
                    if (accept(T_ELSEIF)) {
                        s->conds.push_back(expr());
                        expect(T_THEN, "then");
                        s->blocks.push_back(block());
                    } else if (accept(T_ELSE)) {
                        s->else_body = block();
                        expect(T_END, "end");
                        break;
                    } else {
                        expect(T_END, "end");
                        break;
                    }
                }
                return s;
            }
            case T_FOR: {
                ++i_;
                const std::string n1 = name();
                if (accept(T_ASSIGN)) {
                    auto s = std::make_unique<Stat>(Stat::NUMFOR);
                    s->line = line;
                    s->names.push_back(n1);
                    s->exprs.push_back(expr());
                    expect(T_COMMA, ",");
                    s->exprs.push_back(expr());
                    if (accept(T_COMMA)) s->exprs.push_back(expr());
                    expect(T_DO, "do");
                    s->body = block();
                    expect(T_END, "end");
                    return s;
                }
                auto s = std::make_unique<Stat>(Stat::GENFOR);
                s->line = line;
                s->names.push_back(n1);
                while (accept(T_COMMA)) s->names.push_back(name());
                expect(T_IN, "in");
                s->exprs = exprlist();
                expect(T_DO, "do");
                s->body = block();
                expect(T_END, "end");
                return s;
            }
            case T_FUNCTION: {
                ++i_;
                // funcname: Name {'.' Name} [':' Name]
                auto target = std::make_unique<Expr>(Expr::NAME);
                target->line = line;
                target->s = name();
                std::string full = target->s;
                bool method = false;
                while (cur().t == T_DOT || cur().t == T_COLON) {
                    method = cur().t == T_COLON;
                    ++i_;
                    auto idx = std::make_unique<Expr>(Expr::INDEX);
                    idx->line = line;
                    idx->a = std::move(target);
                    idx->b = std::make_unique<Expr>(Expr::STR);
                    idx->b->s = name();
                    full += (method ? ":" : ".") + idx->b->s;
                    target = std::move(idx);
                    if (method) break;
                }
                auto s = std::make_unique<Stat>(Stat::FUNC);
                s->line = line;
                s->targets.push_back(std::move(target));
                s->fn = funcbody(full, method);
                return s;
            }
            case T_LOCAL: {
                ++i_;
                if (accept(T_FUNCTION)) {
                    auto s = std::make_unique<Stat>(Stat::LOCALFUNC);
                    s->line = line;
                    s->names.push_back(name());
                    s->fn = funcbody(s->names[0], false);
                    return s;
                }
                auto s = std::make_unique<Stat>(Stat::LOCAL);
                s->line = line;
                do {
                    s->names.push_back(name());
                    if (accept(T_LT)) {  // <const> / <close> attributes
                        name();
                        expect(T_GT, ">");
                    }
                } while (accept(T_COMMA));
                if (accept(T_ASSIGN)) s->exprs = exprlist();
                return s;
            }
            case T_GOTO: fail("goto is not supported");
            default: break;
        }
        // exprstat: call or assignment
        ExprP e = suffixedexp();
        if (cur().t == T_ASSIGN || cur().t == T_COMMA) {
            auto s = std::make_unique<Stat>(Stat::ASSIGN);
            s->line = line;
            if (e->k != Expr::NAME && e->k != Expr::INDEX) fail("syntax error");
            s->targets.push_back(std::move(e));
            while (accept(T_COMMA)) {
                ExprP t = suffixedexp();
                if (t->k != Expr::NAME && t->k != Expr::INDEX) fail("syntax error");
                s->targets.push_back(std::move(t));
            }
            expect(T_ASSIGN, "=");
            s->exprs = exprlist();
            return s;
        }
        if (e->k != Expr::CALL && e->k != Expr::METHOD) fail("syntax error");
        auto s = std::make_unique<Stat>(Stat::CALL);
        s->line = line;
        s->exprs.push_back(std::move(e));
        return s;
    }
    std::vector<ExprP> exprlist() {
        std::vector<ExprP> v;
        v.push_back(expr());
        while (accept(T_COMMA)) v.push_back(expr());
        return v;
    }
    ExprP primaryexp() {
        const int line = cur().line;
        if (cur().t == T_NAME) {
            auto e = std::make_unique<Expr>(Expr::NAME);
            e->line = line;
            e->s = name();
            return e;
        }
        if (accept(T_LPAREN)) {
            auto e = std::make_unique<Expr>(Expr::PAREN);
            e->line = line;
            e->a = expr();
            expect(T_RPAREN, ")");
            return e;
        }
        fail("unexpected symbol");
    }
    std::vector<ExprP> callargs() {
        std::vector<ExprP> args;
        if (cur().t == T_STRING) {
            auto e = std::make_unique<Expr>(Expr::STR);
            e->line = cur().line;
            e->s = t_[i_++].s;
            args.push_back(std::move(e));
        } else if (cur().t == T_LBRACE) {
            args.push_back(tablecons());
        } else {
            expect(T_LPAREN, "(");
            if (cur().t != T_RPAREN) args = exprlist();
            expect(T_RPAREN, ")");
        }
        return args;
    }
    ExprP suffixedexp() {
        ExprP e = primaryexp();
        for (;;) {
            const int line = cur().line;
            switch (cur().t) {
                case T_DOT: {
                    ++i_;
                    auto x = std::make_unique<Expr>(Expr::INDEX);
                    x->line = line;
                    x->a = std::move(e);
                    x->b = std::make_unique<Expr>(Expr::STR);
                    x->b->s = name();
                    e = std::move(x);
                    break;
                }
                case T_LBRACKET: {
                    ++i_;
                    auto x = std::make_unique<Expr>(Expr::INDEX);
                    x->line = line;
                    x->a = std::move(e);
                    x->b = expr();
                    expect(T_RBRACKET, "]");
                    e = std::move(x);
                    break;
                }
                case T_COLON: {
                    ++i_;
                    auto x = std::make_unique<Expr>(Expr::METHOD);
                    x->line = line;
                    x->a = std::move(e);
                    x->s = name();
                    x->args = callargs();
                    e = std::move(x);
                    break;
                }
                case T_LPAREN:
                case T_STRING:
                case T_LBRACE: {
                    auto x = std::make_unique<Expr>(Expr::CALL);
                    x->line = line;
                    x->a = std::move(e);
                    x->args = callargs();
                    e = std::move(x);
                    break;
                }
                default: return e;
            }
        }
    }
    ExprP tablecons() {
        auto e = std::make_unique<Expr>(Expr::TABLE);
        e->line = cur().line;
        expect(T_LBRACE, "{");
        while (cur().t != T_RBRACE) {
            TableField f;
            if (cur().t == T_LBRACKET) {
                ++i_;
                f.kind = 2;
                f.key = expr();
                expect(T_RBRACKET, "]");
                expect(T_ASSIGN, "=");
                f.val = expr();
            } else if (cur().t == T_NAME && look(1).t == T_ASSIGN) {
                f.kind = 1;
                f.name = name();
                ++i_;
                f.val = expr();
            } else {
                f.kind = 0;
                f.val = expr();
            }
            e->fields.push_back(std::move(f));
            if (!accept(T_COMMA) && !accept(T_SEMI)) break;
        }
        expect(T_RBRACE, "}");
        return e;
    }
    ExprP simpleexp() {
        const int line = cur().line;
        switch (cur().t) {
            case T_NUMBER: {
                auto e = std::make_unique<Expr>(Expr::NUM);
                e->line = line;
                e->isint = t_[i_].isint;
                e->n = t_[i_++].n;
                return e;
            }
            case T_STRING: { auto e = std::make_unique<Expr>(Expr::STR); e->line = line; e->s = t_[i_++].s; return e; }
            case T_NIL: { ++i_; auto e = std::make_unique<Expr>(Expr::NIL); e->line = line; return e; }
            case T_TRUE: { ++i_; auto e = std::make_unique<Expr>(Expr::TRUE_); e->line = line; return e; }
            case T_FALSE: { ++i_; auto e = std::make_unique<Expr>(Expr::FALSE_); e->line = line; return e; }
            case T_ELLIPSIS: { ++i_; auto e = std::make_unique<Expr>(Expr::VARARG); e->line = line; return e; }
            case T_LBRACE: return tablecons();
            case T_FUNCTION: {
                ++i_;
                auto e = std::make_unique<Expr>(Expr::FUNC);
                e->line = line;
                e->fn = funcbody("anonymous", false);
                return e;
            }
            default: return suffixedexp();
        }
    }
    // binary priorities (left, right) as in Lua 5.3's parser
    static bool binop(Tok t, int& l, int& r) {
        switch (t) {
            case T_OR: l = 1; r = 1; return true;
            case T_AND: l = 2; r = 2; return true;
            case T_LT: case T_GT: case T_LE: case T_GE: case T_NE: case T_EQ: l = 3; r = 3; return true;
            case T_CONCAT: l = 9; r = 8; return true;  // right associative
            case T_PLUS: case T_MINUS: l = 10; r = 10; return true;
            case T_STAR: case T_SLASH: case T_DSLASH: case T_PERCENT: l = 11; r = 11; return true;
            case T_CARET: l = 14; r = 13; return true;  // right associative
            default: return false;
        }
    }
    static constexpr int kUnaryPriority = 12;
    ExprP subexpr(int limit) {
        ExprP e;
        const Tok t = cur().t;
        if (t == T_NOT || t == T_MINUS || t == T_HASH) {
            const int line = cur().line;
            ++i_;
            e = std::make_unique<Expr>(Expr::UN);
            e->line = line;
            e->op = t;
            e->a = subexpr(kUnaryPriority);
        } else {
            e = simpleexp();
        }
        int l, r;
        while (binop(cur().t, l, r) && l > limit) {
            const Tok op = cur().t;
            const int line = cur().line;
            ++i_;
            ExprP rhs = subexpr(r);
            auto x = std::make_unique<Expr>(op == T_AND ? Expr::AND : op == T_OR ? Expr::OR : Expr::BIN);
            x->line = line;
            x->op = op;
            x->a = std::move(e);
            x->b = std::move(rhs);
            e = std::move(x);
        }
        return e;
    }
    ExprP expr() { return subexpr(0); }

    std::vector<Token> t_;
    std::string chunk_;
    size_t i_ = 0;
};

// ===========================================================================
// Interpreter
// ===========================================================================

struct Cell {
    Value v;
};
using CellP = std::shared_ptr<Cell>;

struct Env {
    std::vector<std::pair<std::string, CellP>> vars;
    std::shared_ptr<Env> parent;
    CellP find(const std::string& n) const {
        for (const Env* e = this; e; e = e->parent.get())
            for (auto it = e->vars.rbegin(); it != e->vars.rend(); ++it)
                if (it->first == n) return it->second;
        return nullptr;
    }
};
using EnvP = std::shared_ptr<Env>;

using Native = std::function<Values(Interp&, Values&)>;

struct Function {
    std::shared_ptr<FuncBody> body;  // Lua function
    EnvP env;                        // its enclosing scope
    std::string chunk;
    Native native;                   // or a built-in
    std::string native_name;
};

struct ExitRequest {};

struct Frame {
    Values varargs;
    std::string chunk;
};

struct Interp {
    std::shared_ptr<Table> G;
    std::shared_ptr<Table> loaded;  // package.loaded
    std::string workdir;
    std::string printed;
    std::string last_line;
    GlibcRandom rng;
    std::vector<std::shared_ptr<FuncBody>> keep;  // chunks stay alive while closures refer to them
    std::function<bool(Interp&, const std::string&)> builtin_module;
    bool builtin_only = false;  // lua_core modules always come from builtin_module
    int depth = 0;
    std::vector<Frame*> frames;

    enum Flow { NORMAL, BREAK, RETURN };
    Values ret;

    [[noreturn]] void fail(int line, const std::string& msg) const {
        std::string where = frames.empty() ? std::string("?") : frames.back()->chunk;
        throw std::runtime_error(where + ":" + std::to_string(line) + ": " + msg);
    }

    // --- helpers -----------------------------------------------------------
    static Value make_table() {
        Value v;
        v.k = Value::TABLE;
        v.t = std::make_shared<Table>();
        return v;
    }
    static Value make_native(const std::string& name, Native fn) {
        Value v;
        v.k = Value::FUNC;
        v.f = std::make_shared<Function>();
        v.f->native = std::move(fn);
        v.f->native_name = name;
        return v;
    }
    // lua_Number / lua_Integer formatting: "%d" for integers, "%.14g" for
    // floats with ".0" added when the text looks like an integer
    static std::string num_to_string(const Value& v) {
        const double d = v.n;
        char buf[64];
        if (v.isint) {
            std::snprintf(buf, sizeof buf, "%lld", (long long)d);
            return buf;
        }
        if (std::isinf(d)) return d > 0 ? "inf" : "-inf";
        if (std::isnan(d)) return "nan";
        std::snprintf(buf, sizeof buf, "%.14g", d);
        std::string out = buf;
        if (out.find_first_of(".eEn") == std::string::npos) out += ".0";
        return out;
    }
    std::string tostring(const Value& v) const {
        switch (v.k) {
            case Value::NIL: return "nil";
            case Value::BOOL: return v.b ? "true" : "false";
            case Value::NUM: return num_to_string(v);
            case Value::STR: return *v.s;
            case Value::TABLE: {
                char buf[48];
                std::snprintf(buf, sizeof buf, "table: %p", (const void*)v.t.get());
                return buf;
            }
            default: {
                char buf[48];
                std::snprintf(buf, sizeof buf, "function: %p", (const void*)v.f.get());
                return buf;
            }
        }
    }
    // arithmetic operand: a number, or a string that converts to one
    bool arith_operand(const Value& v, Value& out) const {
        if (v.k == Value::NUM) {
            out = v;
            return true;
        }
        if (v.k == Value::STR) {
            out = str_to_number(*v.s);
            return out.k == Value::NUM;
        }
        return false;
    }
    // tonumber's conversion: integer for an integer literal, float otherwise
    static Value str_to_number(const std::string& str) {
        const char* s = str.c_str();
        char* end = nullptr;
        while (std::isspace((unsigned char)*s)) ++s;
        if (!*s) return Value();
        const char* t = s;
        if (*t == '-' || *t == '+') ++t;
        if (t[0] == '0' && (t[1] == 'x' || t[1] == 'X')) {
            const unsigned long long u = std::strtoull(t + 2, &end, 16);
            while (end && std::isspace((unsigned char)*end)) ++end;
            if (!end || *end || end == t + 2) return Value();
            return Value::integer((double)(int64_t)(s[0] == '-' ? 0 - u : u));
        }
        const double d = std::strtod(s, &end);
        while (end && std::isspace((unsigned char)*end)) ++end;
        if (!end || *end) return Value();
        const bool isint = str.find_first_of(".eEnN") == std::string::npos && std::fabs(d) < 9.2e18;
        return isint ? Value::integer(d) : Value::number(d);
    }
    // string -> number coercion of arithmetic (Lua converts numeric strings)
    bool to_number(const Value& v, double& out) const {
        if (v.k == Value::NUM) {
            out = v.n;
            return true;
        }
        if (v.k == Value::STR) {
            const char* s = v.s->c_str();
            char* end = nullptr;
            while (std::isspace((unsigned char)*s)) ++s;
            if (!*s) return false;
            out = (s[0] == '0' && (s[1] == 'x' || s[1] == 'X')) ? (double)std::strtoull(s + 2, &end, 16)
                                                                  : std::strtod(s, &end);
            while (end && std::isspace((unsigned char)*end)) ++end;
            return end && !*end;
        }
        return false;
    }

    Value index(const Value& obj, const Value& key, int line) {
        if (obj.k == Value::TABLE) return obj.t->get(key);
        if (obj.k == Value::STR) {  // string methods: ("x"):upper()
            Value strlib = G->get("string");
            if (strlib.k == Value::TABLE) return strlib.t->get(key);
        }
        fail(line, std::string("attempt to index a ") + type_name(obj) + " value" +
                       (key.k == Value::STR ? " (field '" + *key.s + "')" : ""));
    }
    void setindex(const Value& obj, const Value& key, const Value& val, int line) {
        if (obj.k != Value::TABLE) fail(line, std::string("attempt to index a ") + type_name(obj) + " value");
        if (key.k == Value::NIL) fail(line, "table index is nil");
        if (key.k == Value::NUM && std::isnan(key.n)) fail(line, "table index is NaN");
        obj.t->set(key, val);
    }

    Values call(const Value& fn, Values& args, int line) {
        if (fn.k != Value::FUNC) fail(line, std::string("attempt to call a ") + type_name(fn) + " value");
        if (++depth > 200) fail(line, "stack overflow");
        struct Guard {
            int& d;
            ~Guard() { --d; }
        } guard{depth};
        const Function& f = *fn.f;
        if (f.native) return f.native(*this, args);
        auto env = std::make_shared<Env>();
        env->parent = f.env;
        for (size_t i = 0; i < f.body->params.size(); ++i)
            env->vars.emplace_back(f.body->params[i],
                                   std::make_shared<Cell>(Cell{i < args.size() ? args[i] : Value()}));
        Frame fr;
        fr.chunk = f.chunk;
        if (f.body->vararg && args.size() > f.body->params.size())
            fr.varargs.assign(args.begin() + (long)f.body->params.size(), args.end());
        frames.push_back(&fr);
        struct Pop {
            std::vector<Frame*>& v;
            ~Pop() { v.pop_back(); }
        } pop{frames};
        const Flow fl = exec_block(f.body->body, env);
        if (fl == RETURN) {
            Values r;
            r.swap(ret);
            return r;
        }
        return {};
    }

    // --- expressions ---------------------------------------------------------
    Value eval1(const Expr& e, const EnvP& env) {
        if (e.multi()) {
            Values v = evalmulti(e, env);
            return v.empty() ? Value() : v[0];
        }
        switch (e.k) {
            case Expr::NIL: return Value();
            case Expr::TRUE_: return Value::boolean(true);
            case Expr::FALSE_: return Value::boolean(false);
            case Expr::NUM: return e.isint ? Value::integer(e.n) : Value::number(e.n);
            case Expr::STR: return Value::string(e.s);
            case Expr::PAREN: return eval1(*e.a, env);
            case Expr::FUNC: {
                Value v;
                v.k = Value::FUNC;
                v.f = std::make_shared<Function>();
                v.f->body = e.fn;
                v.f->env = env;
                v.f->chunk = frames.empty() ? std::string("?") : frames.back()->chunk;
                return v;
            }
            case Expr::TABLE: {
                Value t = make_table();
                double pos = 1.0;
                for (size_t i = 0; i < e.fields.size(); ++i) {
                    const TableField& f = e.fields[i];
                    if (f.kind == 0) {
                        if (i + 1 == e.fields.size() && f.val->multi()) {
                            for (const Value& v : evalmulti(*f.val, env)) t.t->set(Value::integer(pos++), v);
                        } else {
                            Value v = eval1(*f.val, env);
                            t.t->set(Value::integer(pos++), v);
                        }
                    } else if (f.kind == 1) {
                        Value v = eval1(*f.val, env);
                        t.t->set(f.name, v);
                    } else {
                        Value k = eval1(*f.key, env);
                        if (k.k == Value::NIL) fail(e.line, "table index is nil");
                        Value v = eval1(*f.val, env);
                        t.t->set(k, v);
                    }
                }
                return t;
            }
            case Expr::NAME: {
                CellP c = env->find(e.s);
                return c ? c->v : G->get(e.s);
            }
            case Expr::INDEX: {
                Value o = eval1(*e.a, env);
                Value k = eval1(*e.b, env);
                return index(o, k, e.line);
            }
            case Expr::AND: {
                Value a = eval1(*e.a, env);
                return a.truthy() ? eval1(*e.b, env) : a;
            }
            case Expr::OR: {
                Value a = eval1(*e.a, env);
                return a.truthy() ? a : eval1(*e.b, env);
            }
            case Expr::UN: {
                Value a = eval1(*e.a, env);
                if (e.op == T_NOT) return Value::boolean(!a.truthy());
                if (e.op == T_HASH) {
                    if (a.k == Value::STR) return Value::integer((double)a.s->size());
                    if (a.k == Value::TABLE) return Value::integer((double)a.t->length());
                    fail(e.line, std::string("attempt to get length of a ") + type_name(a) + " value");
                }
                Value x;
                if (!arith_operand(a, x)) fail(e.line, std::string("attempt to perform arithmetic on a ") + type_name(a) + " value");
                return x.isint ? Value::integer(-x.n) : Value::number(-x.n);
            }
            case Expr::BIN: {
                Value a = eval1(*e.a, env);
                Value b = eval1(*e.b, env);
                return binop(e.op, a, b, e.line);
            }
            default: fail(e.line, "bad expression");
        }
    }
    Value binop(Tok op, const Value& a, const Value& b, int line) {
        switch (op) {
            case T_EQ: return Value::boolean(raw_equal(a, b));
            case T_NE: return Value::boolean(!raw_equal(a, b));
            case T_LT: case T_LE: case T_GT: case T_GE: {
                if (a.k == Value::NUM && b.k == Value::NUM) {
                    const double x = a.n, y = b.n;
                    return Value::boolean(op == T_LT ? x < y : op == T_LE ? x <= y : op == T_GT ? x > y : x >= y);
                }
                if (a.k == Value::STR && b.k == Value::STR) {
                    const int c = a.s->compare(*b.s);
                    return Value::boolean(op == T_LT ? c < 0 : op == T_LE ? c <= 0 : op == T_GT ? c > 0 : c >= 0);
                }
                fail(line, std::string("attempt to compare ") + type_name(a) + " with " + type_name(b));
            }
            case T_CONCAT: {
                auto piece = [&](const Value& v) -> std::string {
                    if (v.k == Value::STR) return *v.s;
                    if (v.k == Value::NUM) return num_to_string(v);
                    fail(line, std::string("attempt to concatenate a ") + type_name(v) + " value");
                };
                return Value::string(piece(a) + piece(b));
            }
            default: break;
        }
        Value xa, yb;
        if (!arith_operand(a, xa) || !arith_operand(b, yb))
            fail(line, std::string("attempt to perform arithmetic on a ") + type_name(arith_operand(a, xa) ? b : a) +
                           " value");
        const double x = xa.n, y = yb.n;
        const bool ints = xa.isint && yb.isint;  // integer op integer stays integer, except / and ^
        switch (op) {
            case T_PLUS: return ints ? Value::integer(x + y) : Value::number(x + y);
            case T_MINUS: return ints ? Value::integer(x - y) : Value::number(x - y);
            case T_STAR: return ints ? Value::integer(x * y) : Value::number(x * y);
            case T_SLASH: return Value::number(x / y);
            case T_DSLASH:
                if (ints && y == 0.0) fail(line, "attempt to perform 'n//0'");
                return ints ? Value::integer(std::floor(x / y)) : Value::number(std::floor(x / y));
            case T_PERCENT: {
                if (ints && y == 0.0) fail(line, "attempt to perform 'n%%0'");
                // a - floor(a / b) * b: fmod with the sign of the divisor
                double m = std::fmod(x, y);
                if (m != 0.0 && ((m < 0) != (y < 0))) m += y;
                return ints ? Value::integer(m) : Value::number(m);
            }
            case T_CARET: return Value::number(std::pow(x, y));
            default: fail(line, "bad operator");
        }
    }
    Values evalmulti(const Expr& e, const EnvP& env) {
        if (e.k == Expr::VARARG) {
            if (frames.empty()) return {};
            return frames.back()->varargs;
        }
        if (e.k == Expr::CALL) {
            Value fn = eval1(*e.a, env);
            Values args = evalargs(e.args, env);
            return call(fn, args, e.line);
        }
        if (e.k == Expr::METHOD) {
            Value obj = eval1(*e.a, env);
            Value fn = index(obj, Value::string(e.s), e.line);
            Values args = evalargs(e.args, env);
            args.insert(args.begin(), obj);
            return call(fn, args, e.line);
        }
        return {eval1(e, env)};
    }
    Values evalargs(const std::vector<ExprP>& list, const EnvP& env) {
        Values out;
        for (size_t i = 0; i < list.size(); ++i) {
            if (i + 1 == list.size() && list[i]->multi()) {
                Values m = evalmulti(*list[i], env);
                out.insert(out.end(), m.begin(), m.end());
            } else {
                out.push_back(eval1(*list[i], env));
            }
        }
        return out;
    }

    // --- statements -----------------------------------------------------------
    void assign(const Expr& target, const Value& v, const EnvP& env) {
        if (target.k == Expr::NAME) {
            CellP c = env->find(target.s);
            if (c)
                c->v = v;
            else
                G->set(target.s, v);
            return;
        }
        Value o = eval1(*target.a, env);
        Value k = eval1(*target.b, env);
        setindex(o, k, v, target.line);
    }
    Flow exec_block(const Block& b, const EnvP& parent) {
        auto env = std::make_shared<Env>();
        env->parent = parent;
        for (const StatP& s : b) {
            const Flow f = exec(*s, env);
            if (f != NORMAL) return f;
        }
        return NORMAL;
    }
    Flow exec(const Stat& s, const EnvP& env) {
        switch (s.k) {
            case Stat::LOCAL: {
                Values v = evalargs(s.exprs, env);
                for (size_t i = 0; i < s.names.size(); ++i)
                    env->vars.emplace_back(s.names[i], std::make_shared<Cell>(Cell{i < v.size() ? v[i] : Value()}));
                return NORMAL;
            }
            case Stat::ASSIGN: {
                Values v = evalargs(s.exprs, env);
                for (size_t i = 0; i < s.targets.size(); ++i) assign(*s.targets[i], i < v.size() ? v[i] : Value(), env);
                return NORMAL;
            }
            case Stat::CALL: evalmulti(*s.exprs[0], env); return NORMAL;
            case Stat::DO: return exec_block(s.body, env);
            case Stat::WHILE:
                while (eval1(*s.cond, env).truthy()) {
                    const Flow f = exec_block(s.body, env);
                    if (f == BREAK) break;
                    if (f == RETURN) return f;
                }
                return NORMAL;
            case Stat::REPEAT:
                for (;;) {
                    // the condition sees the body's locals
                    auto inner = std::make_shared<Env>();
                    inner->parent = env;
                    Flow f = NORMAL;
                    for (const StatP& st : s.body) {
                        f = exec(*st, inner);
                        if (f != NORMAL) break;
                    }
                    if (f == BREAK) break;
                    if (f == RETURN) return f;
                    if (eval1(*s.cond, inner).truthy()) break;
                }
                return NORMAL;
            case Stat::IF:
                for (size_t i = 0; i < s.conds.size(); ++i)
                    if (eval1(*s.conds[i], env).truthy()) return exec_block(s.blocks[i], env);
                return exec_block(s.else_body, env);
            case Stat::NUMFOR: {
                const Value va = eval1(*s.exprs[0], env), vb = eval1(*s.exprs[1], env);
                const Value vc = s.exprs.size() > 2 ? eval1(*s.exprs[2], env) : Value::integer(1);
                if (va.k != Value::NUM) fail(s.line, "'for' initial value must be a number");
                if (vb.k != Value::NUM) fail(s.line, "'for' limit must be a number");
                if (vc.k != Value::NUM) fail(s.line, "'for' step must be a number");
                const double a = va.n, c = vc.n;
                if (c == 0.0) fail(s.line, "'for' step is zero");
                // integer loop when start and step are integers (the limit is floored / ceiled)
                const bool ints = va.isint && vc.isint;
                const double b = ints ? (c > 0 ? std::floor(vb.n) : std::ceil(vb.n)) : vb.n;
                for (double i = a; c > 0 ? i <= b : i >= b; i += c) {
                    auto loop = std::make_shared<Env>();
                    loop->parent = env;
                    loop->vars.emplace_back(s.names[0],
                                            std::make_shared<Cell>(Cell{ints ? Value::integer(i) : Value::number(i)}));
                    const Flow f = exec_block(s.body, loop);
                    if (f == BREAK) break;
                    if (f == RETURN) return f;
                }
                return NORMAL;
            }
            case Stat::GENFOR: {
                Values init = evalargs(s.exprs, env);
                init.resize(3);
                Value fn = init[0], state = init[1], ctl = init[2];
                for (;;) {
                    Values args{state, ctl};
                    Values r = call(fn, args, s.line);
                    r.resize(std::max<size_t>(r.size(), s.names.size()));
                    if (r[0].is_nil()) break;
                    ctl = r[0];
                    auto loop = std::make_shared<Env>();
                    loop->parent = env;
                    for (size_t i = 0; i < s.names.size(); ++i)
                        loop->vars.emplace_back(s.names[i], std::make_shared<Cell>(Cell{r[i]}));
                    const Flow f = exec_block(s.body, loop);
                    if (f == BREAK) break;
                    if (f == RETURN) return f;
                }
                return NORMAL;
            }
            case Stat::FUNC: {
                Value v;
                v.k = Value::FUNC;
                v.f = std::make_shared<Function>();
                v.f->body = s.fn;
                v.f->env = env;
                v.f->chunk = frames.empty() ? std::string("?") : frames.back()->chunk;
                assign(*s.targets[0], v, env);
                return NORMAL;
            }
            case Stat::LOCALFUNC: {
                auto cell = std::make_shared<Cell>();
                env->vars.emplace_back(s.names[0], cell);  // visible inside its own body (recursion)
                Value v;
                v.k = Value::FUNC;
                v.f = std::make_shared<Function>();
                v.f->body = s.fn;
                v.f->env = env;
                v.f->chunk = frames.empty() ? std::string("?") : frames.back()->chunk;
                cell->v = v;
                return NORMAL;
            }
            case Stat::RETURN: {
                if (s.exprs.size() == 1 && s.exprs[0]->multi()) {
                    ret = evalmulti(*s.exprs[0], env);  // tail call: all its values
                } else {
                    ret = evalargs(s.exprs, env);
                }
                return RETURN;
            }
            case Stat::BREAK: return BREAK;
        }
        return NORMAL;
    }

    // Runs a source file as a chunk (a function with varargs).
    Values run_file(const std::string& path, const std::string& chunkname, Values args) {
        std::ifstream in(path, std::ios::binary);
        if (!in) throw std::runtime_error("cannot open " + path);
        std::stringstream ss;
        ss << in.rdbuf();
        std::string src = ss.str();
        if (src.size() >= 1 && src[0] == '#') {  // shebang line
            const size_t e = src.find('\n');
            src = e == std::string::npos ? std::string() : std::string(e, ' ') + src.substr(e);
        }
        Lexer lx(src, chunkname);
        Parser ps(lx.run(), chunkname);
        std::shared_ptr<FuncBody> body = ps.chunk();
        keep.push_back(body);
        Value fn;
        fn.k = Value::FUNC;
        fn.f = std::make_shared<Function>();
        fn.f->body = body;
        fn.f->env = nullptr;
        fn.f->chunk = chunkname;
        return call(fn, args, 0);
    }
};

// dir/p with "." and ".." folded lexically, so a working directory that is
// only named (the reference's build/ directory, say) need not exist
std::string join_path(const std::string& dir, const std::string& p) {
    if (p.empty() || p[0] == '/' || dir.empty()) return p;
    return (std::filesystem::path(dir) / p).lexically_normal().string();
}

bool file_readable(const std::string& p) {
    std::ifstream f(p);
    return (bool)f;
}

// ===========================================================================
// Standard library subset
// ===========================================================================

double check_num(Interp& I, Values& a, size_t i, const char* fname) {
    double d;
    if (i >= a.size() || !I.to_number(a[i], d))
        throw std::runtime_error(std::string("bad argument #") + std::to_string(i + 1) + " to '" + fname +
                                 "' (number expected, got " + (i < a.size() ? type_name(a[i]) : "no value") + ")");
    return d;
}
std::string check_str(Interp& I, Values& a, size_t i, const char* fname) {
    if (i < a.size() && a[i].k == Value::STR) return *a[i].s;
    if (i < a.size() && a[i].k == Value::NUM) return I.tostring(a[i]);
    throw std::runtime_error(std::string("bad argument #") + std::to_string(i + 1) + " to '" + fname +
                             "' (string expected, got " + (i < a.size() ? type_name(a[i]) : "no value") + ")");
}
std::shared_ptr<Table> check_table(Values& a, size_t i, const char* fname) {
    if (i < a.size() && a[i].k == Value::TABLE) return a[i].t;
    throw std::runtime_error(std::string("bad argument #") + std::to_string(i + 1) + " to '" + fname +
                             "' (table expected, got " + (i < a.size() ? type_name(a[i]) : "no value") + ")");
}

// string.format: %d %i %u %c %x %X %o %e %E %f %g %G %q %s %a %%
std::string lua_format(Interp& I, Values& a) {
    const std::string fmt = check_str(I, a, 0, "format");
    std::string out;
    size_t arg = 1;
    for (size_t i = 0; i < fmt.size(); ++i) {
        if (fmt[i] != '%') {
            out += fmt[i];
            continue;
        }
        if (++i >= fmt.size()) throw std::runtime_error("invalid conversion '%' to 'format'");
        if (fmt[i] == '%') {
            out += '%';
            continue;
        }
        std::string spec = "%";
        while (i < fmt.size() && std::strchr("-+ #0123456789.", fmt[i])) spec += fmt[i++];
        if (i >= fmt.size()) throw std::runtime_error("invalid conversion to 'format'");
        const char c = fmt[i];
        char buf[512];
        switch (c) {
            case 'd': case 'i': {
                spec += "lld";
                std::snprintf(buf, sizeof buf, spec.c_str(), (long long)check_num(I, a, arg++, "format"));
                out += buf;
                break;
            }
            case 'u': case 'x': case 'X': case 'o': case 'c': {
                spec += (c == 'c') ? std::string("c") : std::string("ll") + c;
                if (c == 'c')
                    std::snprintf(buf, sizeof buf, spec.c_str(), (int)check_num(I, a, arg++, "format"));
                else
                    std::snprintf(buf, sizeof buf, spec.c_str(), (unsigned long long)(long long)check_num(I, a, arg++, "format"));
                out += buf;
                break;
            }
            case 'e': case 'E': case 'f': case 'F': case 'g': case 'G': case 'a': case 'A': {
                spec += c;
                std::snprintf(buf, sizeof buf, spec.c_str(), check_num(I, a, arg++, "format"));
                out += buf;
                break;
            }
            case 's': {
                spec += 's';
                const std::string s = arg < a.size() ? I.tostring(a[arg]) : std::string("nil");
                ++arg;
                if (spec == "%s") {
                    out += s;
                } else {
                    std::snprintf(buf, sizeof buf, spec.c_str(), s.c_str());
                    out += buf;
                }
                break;
            }
            case 'q': {
                out += '"' + check_str(I, a, arg++, "format") + '"';
                break;
            }
            default: throw std::runtime_error(std::string("invalid conversion '%") + c + "' to 'format'");
        }
    }
    return out;
}

void open_libs(Interp& I) {
    Table& G = *I.G;
    Value gv;
    gv.k = Value::TABLE;
    gv.t = I.G;
    G.set("_G", gv);
    G.set("_VERSION", Value::string("Lua 5.3"));
    G.set("print", Interp::make_native("print", [](Interp& I, Values& a) {
        std::string line;
        for (size_t i = 0; i < a.size(); ++i) line += (i ? "\t" : "") + I.tostring(a[i]);
        I.printed += line + "\n";
        I.last_line = line;
        return Values{};
    }));
    G.set("type", Interp::make_native("type", [](Interp&, Values& a) {
        if (a.empty()) throw std::runtime_error("bad argument #1 to 'type' (value expected)");
        return Values{Value::string(type_name(a[0]))};
    }));
    G.set("tostring", Interp::make_native("tostring", [](Interp& I, Values& a) {
        return Values{Value::string(I.tostring(a.empty() ? Value() : a[0]))};
    }));
    G.set("tonumber", Interp::make_native("tonumber", [](Interp&, Values& a) {
        if (!a.empty() && a.size() > 1 && a[1].k == Value::NUM && a[0].k == Value::STR) {
            char* end = nullptr;
            const long long v = std::strtoll(a[0].s->c_str(), &end, (int)a[1].n);
            return Values{(end && !*end) ? Value::integer((double)v) : Value()};
        }
        if (!a.empty() && a[0].k == Value::NUM) return Values{a[0]};
        if (!a.empty() && a[0].k == Value::STR) return Values{Interp::str_to_number(*a[0].s)};
        return Values{Value()};
    }));
    G.set("rawget", Interp::make_native("rawget", [](Interp&, Values& a) {
        auto t = check_table(a, 0, "rawget");
        return Values{t->get(a.size() > 1 ? a[1] : Value())};
    }));
    G.set("rawset", Interp::make_native("rawset", [](Interp&, Values& a) {
        auto t = check_table(a, 0, "rawset");
        if (a.size() < 3 || a[1].is_nil()) throw std::runtime_error("bad argument to 'rawset'");
        t->set(a[1], a[2]);
        return Values{a[0]};
    }));
    G.set("rawequal", Interp::make_native("rawequal", [](Interp&, Values& a) {
        a.resize(2);
        return Values{Value::boolean(raw_equal(a[0], a[1]))};
    }));
    G.set("rawlen", Interp::make_native("rawlen", [](Interp&, Values& a) {
        if (!a.empty() && a[0].k == Value::STR) return Values{Value::integer((double)a[0].s->size())};
        return Values{Value::integer((double)check_table(a, 0, "rawlen")->length())};
    }));
    G.set("next", Interp::make_native("next", [](Interp&, Values& a) {
        auto t = check_table(a, 0, "next");
        Value k, v;
        if (!t->next(a.size() > 1 ? a[1] : Value(), k, v)) return Values{Value()};
        return Values{k, v};
    }));
    G.set("pairs", Interp::make_native("pairs", [](Interp& I, Values& a) {
        check_table(a, 0, "pairs");
        return Values{I.G->get("next"), a[0], Value()};
    }));
    G.set("ipairs", Interp::make_native("ipairs", [](Interp&, Values& a) {
        check_table(a, 0, "ipairs");
        Value iter = Interp::make_native("ipairs_iter", [](Interp&, Values& b) {
            const double i = b[1].n + 1.0;
            Value v = b[0].t->get(Value::integer(i));
            if (v.is_nil()) return Values{Value()};
            return Values{Value::integer(i), v};
        });
        return Values{iter, a[0], Value::integer(0.0)};
    }));
    G.set("select", Interp::make_native("select", [](Interp& I, Values& a) {
        if (!a.empty() && a[0].k == Value::STR && *a[0].s == "#") return Values{Value::integer((double)a.size() - 1)};
        const double n = check_num(I, a, 0, "select");
        long long i = (long long)n;
        if (i < 0) i = (long long)a.size() + i;
        else if (i == 0) throw std::runtime_error("bad argument #1 to 'select' (index out of range)");
        if (i >= (long long)a.size()) return Values{};
        return Values(a.begin() + i, a.end());
    }));
    G.set("error", Interp::make_native("error", [](Interp& I, Values& a) -> Values {
        throw std::runtime_error(a.empty() ? std::string("nil") : I.tostring(a[0]));
    }));
    G.set("assert", Interp::make_native("assert", [](Interp& I, Values& a) -> Values {
        if (a.empty() || !a[0].truthy())
            throw std::runtime_error(a.size() > 1 ? I.tostring(a[1]) : std::string("assertion failed!"));
        return a;
    }));
    G.set("unpack", Interp::make_native("unpack", [](Interp&, Values& a) {
        auto t = check_table(a, 0, "unpack");
        Values out;
        for (size_t i = 1; i <= t->length(); ++i) out.push_back(t->get(Value::integer((double)i)));
        return out;
    }));
    G.set("require", Interp::make_native("require", [](Interp& I, Values& a) {
        const std::string name = check_str(I, a, 0, "require");
        Value done = I.loaded->get(name);
        if (!done.is_nil()) return Values{done};
        if (I.builtin_only && I.builtin_module && I.builtin_module(I, name)) {
            I.loaded->set(name, Value::boolean(true));
            return Values{Value::boolean(true)};
        }
        // package.path templates, resolved against the working directory
        Value pkg = I.G->get("package");
        std::string path = pkg.k == Value::TABLE && pkg.t->get("path").k == Value::STR ? *pkg.t->get("path").s : "";
        std::string dotted = name;
        for (char& c : dotted)
            if (c == '.') c = '/';
        std::string tried;
        size_t start = 0;
        while (start <= path.size()) {
            size_t end = path.find(';', start);
            if (end == std::string::npos) end = path.size();
            std::string tmpl = path.substr(start, end - start);
            start = end + 1;
            if (tmpl.empty()) continue;
            std::string file;
            for (char c : tmpl) file += (c == '?') ? dotted : std::string(1, c);
            const std::string full = join_path(I.workdir, file);
            if (file_readable(full)) {
                I.loaded->set(name, Value::boolean(true));  // guards a cyclic require
                Values r = I.run_file(full, file, {Value::string(name), Value::string(file)});
                Value v = r.empty() || r[0].is_nil() ? Value::boolean(true) : r[0];
                I.loaded->set(name, v);
                return Values{v};
            }
            tried += "\n\tno file '" + file + "'";
        }
        if (I.builtin_module && I.builtin_module(I, name)) {
            I.loaded->set(name, Value::boolean(true));
            return Values{Value::boolean(true)};
        }
        throw std::runtime_error("module '" + name + "' not found:" + tried);
    }));

    Value package = Interp::make_table();
    package.t->set("path", Value::string("./?.lua;./?/init.lua"));
    Value loaded;
    loaded.k = Value::TABLE;
    loaded.t = I.loaded;
    package.t->set("loaded", loaded);
    G.set("package", package);

    Value os = Interp::make_table();
    os.t->set("exit", Interp::make_native("exit", [](Interp&, Values&) -> Values { throw ExitRequest{}; }));
    os.t->set("clock", Interp::make_native("clock", [](Interp&, Values&) {
        return Values{Value::number((double)std::clock() / CLOCKS_PER_SEC)};
    }));
    os.t->set("time", Interp::make_native("time", [](Interp&, Values&) {
        return Values{Value::integer((double)std::time(nullptr))};
    }));
    os.t->set("getenv", Interp::make_native("getenv", [](Interp& I, Values& a) {
        const char* v = std::getenv(check_str(I, a, 0, "getenv").c_str());
        return Values{v ? Value::string(v) : Value()};
    }));
    G.set("os", os);

    // math (lmathlib.c of Lua 5.3; random on POSIX = glibc random())
    Value m = Interp::make_table();
    auto fn1 = [](const char* n, double (*f)(double)) {
        return Interp::make_native(n, [n, f](Interp& I, Values& a) { return Values{Value::number(f(check_num(I, a, 0, n)))}; });
    };
    // floor / ceil return integers when the result fits (lmathlib.c math_floor)
    auto fn_int = [](const char* n, double (*f)(double)) {
        return Interp::make_native(n, [n, f](Interp& I, Values& a) {
            if (!a.empty() && a[0].k == Value::NUM && a[0].isint) return Values{a[0]};
            const double r = f(check_num(I, a, 0, n));
            return Values{std::fabs(r) < 9.2e18 ? Value::integer(r) : Value::number(r)};
        });
    };
    m.t->set("floor", fn_int("floor", [](double x) { return std::floor(x); }));
    m.t->set("ceil", fn_int("ceil", [](double x) { return std::ceil(x); }));
    m.t->set("sqrt", fn1("sqrt", [](double x) { return std::sqrt(x); }));
    m.t->set("sin", fn1("sin", [](double x) { return std::sin(x); }));
    m.t->set("cos", fn1("cos", [](double x) { return std::cos(x); }));
    m.t->set("tan", fn1("tan", [](double x) { return std::tan(x); }));
    m.t->set("asin", fn1("asin", [](double x) { return std::asin(x); }));
    m.t->set("acos", fn1("acos", [](double x) { return std::acos(x); }));
    m.t->set("exp", fn1("exp", [](double x) { return std::exp(x); }));
    m.t->set("abs", Interp::make_native("abs", [](Interp& I, Values& a) {
        const double x = std::fabs(check_num(I, a, 0, "abs"));
        return Values{a[0].k == Value::NUM && a[0].isint ? Value::integer(x) : Value::number(x)};
    }));
    m.t->set("rad", fn1("rad", [](double x) { return x * (3.141592653589793 / 180.0); }));
    m.t->set("deg", fn1("deg", [](double x) { return x * (180.0 / 3.141592653589793); }));
    m.t->set("atan", Interp::make_native("atan", [](Interp& I, Values& a) {
        const double y = check_num(I, a, 0, "atan");
        const double x = a.size() > 1 ? check_num(I, a, 1, "atan") : 1.0;
        return Values{Value::number(std::atan2(y, x))};
    }));
    m.t->set("log", Interp::make_native("log", [](Interp& I, Values& a) {
        const double x = check_num(I, a, 0, "log");
        if (a.size() > 1) {
            const double b = check_num(I, a, 1, "log");
            return Values{Value::number(b == 2.0 ? std::log2(x) : b == 10.0 ? std::log10(x) : std::log(x) / std::log(b))};
        }
        return Values{Value::number(std::log(x))};
    }));
    m.t->set("fmod", Interp::make_native("fmod", [](Interp& I, Values& a) {
        return Values{Value::number(std::fmod(check_num(I, a, 0, "fmod"), check_num(I, a, 1, "fmod")))};
    }));
    m.t->set("modf", Interp::make_native("modf", [](Interp& I, Values& a) {
        double ip;
        const double fp = std::modf(check_num(I, a, 0, "modf"), &ip);
        return Values{Value::number(ip), Value::number(fp)};
    }));
    // min / max return the chosen argument itself (its subtype kept)
    m.t->set("min", Interp::make_native("min", [](Interp& I, Values& a) {
        size_t best = 0;
        double r = check_num(I, a, 0, "min");
        for (size_t i = 1; i < a.size(); ++i)
            if (check_num(I, a, i, "min") < r) r = a[i].n, best = i;
        return Values{a[best]};
    }));
    m.t->set("max", Interp::make_native("max", [](Interp& I, Values& a) {
        size_t best = 0;
        double r = check_num(I, a, 0, "max");
        for (size_t i = 1; i < a.size(); ++i)
            if (check_num(I, a, i, "max") > r) r = a[i].n, best = i;
        return Values{a[best]};
    }));
    m.t->set("tointeger", Interp::make_native("tointeger", [](Interp&, Values& a) {
        if (!a.empty() && a[0].k == Value::NUM && std::floor(a[0].n) == a[0].n) return Values{Value::integer(a[0].n)};
        return Values{Value()};
    }));
    m.t->set("type", Interp::make_native("type", [](Interp&, Values& a) {
        if (a.empty() || a[0].k != Value::NUM) return Values{Value()};
        return Values{Value::string(a[0].isint ? "integer" : "float")};
    }));
    m.t->set("huge", Value::number(std::numeric_limits<double>::infinity()));
    m.t->set("pi", Value::number(3.141592653589793));
    m.t->set("maxinteger", Value::integer(9223372036854775807.0));
    m.t->set("mininteger", Value::integer(-9223372036854775808.0));
    // math.random: r = random() / (RAND_MAX + 1); (m, n) -> floor(r * (n - m + 1)) + m
    m.t->set("random", Interp::make_native("random", [](Interp& I, Values& a) {
        const double r = (double)I.rng.next() * (1.0 / (2147483647.0 + 1.0));
        if (a.empty()) return Values{Value::number(r)};
        double lo = 1.0, up;
        if (a.size() == 1) {
            up = check_num(I, a, 0, "random");
        } else {
            lo = check_num(I, a, 0, "random");
            up = check_num(I, a, 1, "random");
        }
        if (!(lo <= up)) throw std::runtime_error("bad argument to 'random' (interval is empty)");
        return Values{Value::integer(std::floor(r * (up - lo + 1.0)) + lo)};
    }));
    // math.randomseed(x): srandom(x), then one draw is discarded
    m.t->set("randomseed", Interp::make_native("randomseed", [](Interp& I, Values& a) {
        I.rng.seed_with((uint32_t)(int64_t)check_num(I, a, 0, "randomseed"));
        (void)I.rng.next();
        return Values{};
    }));
    G.set("math", m);

    Value str = Interp::make_table();
    str.t->set("format", Interp::make_native("format", [](Interp& I, Values& a) {
        return Values{Value::string(lua_format(I, a))};
    }));
    str.t->set("len", Interp::make_native("len", [](Interp& I, Values& a) {
        return Values{Value::integer((double)check_str(I, a, 0, "len").size())};
    }));
    str.t->set("upper", Interp::make_native("upper", [](Interp& I, Values& a) {
        std::string s = check_str(I, a, 0, "upper");
        for (char& c : s) c = (char)std::toupper((unsigned char)c);
        return Values{Value::string(s)};
    }));
    str.t->set("lower", Interp::make_native("lower", [](Interp& I, Values& a) {
        std::string s = check_str(I, a, 0, "lower");
        for (char& c : s) c = (char)std::tolower((unsigned char)c);
        return Values{Value::string(s)};
    }));
    str.t->set("rep", Interp::make_native("rep", [](Interp& I, Values& a) {
        const std::string s = check_str(I, a, 0, "rep");
        const double n = check_num(I, a, 1, "rep");
        std::string out;
        for (long long i = 0; i < (long long)n; ++i) out += s;
        return Values{Value::string(out)};
    }));
    str.t->set("sub", Interp::make_native("sub", [](Interp& I, Values& a) {
        const std::string s = check_str(I, a, 0, "sub");
        const long long len = (long long)s.size();
        long long i = a.size() > 1 ? (long long)check_num(I, a, 1, "sub") : 1;
        long long j = a.size() > 2 ? (long long)check_num(I, a, 2, "sub") : -1;
        if (i < 0) i = std::max(len + i + 1, 1LL);
        else if (i == 0) i = 1;
        if (j < 0) j = len + j + 1;
        else if (j > len) j = len;
        if (i > j) return Values{Value::string("")};
        return Values{Value::string(s.substr((size_t)(i - 1), (size_t)(j - i + 1)))};
    }));
    G.set("string", str);

    Value tab = Interp::make_table();
    tab.t->set("insert", Interp::make_native("insert", [](Interp& I, Values& a) {
        auto t = check_table(a, 0, "insert");
        const size_t n = t->length();
        if (a.size() == 2) {
            t->set(Value::integer((double)n + 1.0), a[1]);
        } else if (a.size() == 3) {
            const double pos = check_num(I, a, 1, "insert");
            if (pos < 1 || pos > (double)n + 1) throw std::runtime_error("bad argument #2 to 'insert' (position out of bounds)");
            for (size_t i = n; i >= (size_t)pos; --i) t->set(Value::integer((double)i + 1.0), t->get(Value::integer((double)i)));
            t->set(Value::integer(pos), a[2]);
        } else {
            throw std::runtime_error("wrong number of arguments to 'insert'");
        }
        return Values{};
    }));
    tab.t->set("concat", Interp::make_native("concat", [](Interp& I, Values& a) {
        auto t = check_table(a, 0, "concat");
        const std::string sep = a.size() > 1 ? check_str(I, a, 1, "concat") : std::string();
        std::string out;
        for (size_t i = 1; i <= t->length(); ++i) {
            Value v = t->get(Value::integer((double)i));
            if (v.k != Value::STR && v.k != Value::NUM) throw std::runtime_error("invalid value in table for 'concat'");
            out += (i > 1 ? sep : std::string()) + I.tostring(v);
        }
        return Values{Value::string(out)};
    }));
    tab.t->set("unpack", I.G->get("unpack"));
    G.set("table", tab);
}

// ===========================================================================
// Built-in scene functions: the behaviour of lua_core/*.lua
// ===========================================================================

struct Core {
    // The registry (lua_core/registervars.lua:1-16).
    static void registervars(Interp& I) {
        Value re = Interp::make_table();
        for (const char* k : {"__CAMERA__", "__PIXEL_SAMPLER__", "__IMAGE_BUFFER__", "__GLOBAL_SETTINGS__", "__OBJECTS__",
                              "__BSDFS__", "__LAYERED_BSDFS__", "__EMISSIONS__", "__MATERIALS__", "__INTEGRATOR__"})
            re.t->set(k, Interp::make_table());
        for (const char* k : {"__OBJECT_COUNT__", "__BSDF_COUNT__", "__LAYERED_BSDF_COUNT__", "__EMISSION_COUNT__",
                              "__MATERIAL_COUNT__"})
            re.t->set(k, Value::integer(0.0));
        I.G->set("__RENDERER_ELEMENTS__", re);
    }
    static Table& reg(Interp& I) {
        Value re = I.G->get("__RENDERER_ELEMENTS__");
        if (re.k != Value::TABLE) throw std::runtime_error("attempt to index a nil value (global '__RENDERER_ELEMENTS__')");
        return *re.t;
    }
    // "Lua ERROR: ... Exiting..." printed, then os.exit()
    [[noreturn]] static void exit_with(Interp& I, const std::string& msg) {
        I.printed += msg + "\n";
        I.last_line = msg;
        throw ExitRequest{};
    }
    static std::shared_ptr<Table> self(Values& a, const char* fname) {
        if (a.empty() || a[0].k != Value::TABLE)
            throw std::runtime_error(std::string("attempt to index a ") + (a.empty() ? "nil" : type_name(a[0])) +
                                     " value (local 'self') in " + fname);
        return a[0].t;
    }
    // validation.lua: CheckVector3 / CheckVector3Field / CheckNumberField / CheckStringField
    static bool vec3(const Value& v) {
        if (v.k != Value::TABLE && v.k != Value::STR)
            throw std::runtime_error(std::string("attempt to get length of a ") + type_name(v) + " value");
        if (v.k == Value::STR) return false;  // #"abc" ~= 3 or v[1] is not a number
        if (v.t->length() != 3) return false;
        for (int i = 1; i <= 3; ++i)
            if (v.t->get(Value::integer(i)).k != Value::NUM) return false;
        return true;
    }
    // a fresh {x, y, z}; components written as integer literals in lua_core stay integers
    static Value vec3_value(std::initializer_list<double> d) {
        Value t = Interp::make_table();
        double i = 1.0;
        for (double x : d) t.t->set(Value::integer(i++), std::floor(x) == x ? Value::integer(x) : Value::number(x));
        return t;
    }
    static bool vec3_field(Table& o, const char* f, const Value& def = Value()) {
        Value v = o.get(f);
        if (v.is_nil()) {
            o.set(f, def);  // a fresh table per call, as the Lua literal is
            return true;
        }
        return vec3(v);
    }
    static bool number_field(Table& o, const char* f, const Value& def = Value()) {
        Value v = o.get(f);
        if (v.is_nil()) {
            if (def.is_nil()) return false;
            o.set(f, def);
            return true;
        }
        return v.k == Value::NUM;
    }
    static bool string_field(Interp& I, Table& o, const char* f, const Value& def = Value(),
                             std::initializer_list<const char*> values = {}) {
        Value v = o.get(f);
        if (v.is_nil()) {
            if (def.is_nil()) return false;
            o.set(f, def);
            return true;
        }
        if (values.size()) {
            bool result = false;
            for (const char* s : values)
                if (v.k == Value::STR && *v.s == s) result = true;
            I.G->set("result", Value::boolean(result));  // validation.lua assigns the global `result`
            return result;
        }
        return v.k == Value::STR;
    }
    static double add_to(Interp& I, const char* list, const char* count, const Value& v) {
        Table& r = reg(I);
        Value c = r.get(count);
        const double n = (c.k == Value::NUM ? c.n : 0.0) + 1.0;
        r.set(count, Value::integer(n));
        Value l = r.get(list);
        if (l.k != Value::TABLE) throw std::runtime_error(std::string("attempt to index a nil value (field '") + list + "')");
        l.t->set(Value::integer(n), v);
        return n;
    }
    static void define(Interp& I, const char* name, std::function<Values(Interp&, Values&)> fn) {
        I.G->set(name, Interp::make_native(name, std::move(fn)));
    }

    static void install(Interp& I) {
        // validation.lua's helpers stay callable from scripts
        define(I, "CheckVector3", [](Interp&, Values& a) { return Values{Value::boolean(vec3(a.empty() ? Value() : a[0]))}; });
        define(I, "CheckVector3Field", [](Interp&, Values& a) {
            auto o = self(a, "CheckVector3Field");
            a.resize(3);
            if (a[1].k != Value::STR) throw std::runtime_error("CheckVector3Field: field name expected");
            return Values{Value::boolean(vec3_field(*o, a[1].s->c_str(), a[2]))};
        });
        define(I, "CheckNumberField", [](Interp&, Values& a) {
            auto o = self(a, "CheckNumberField");
            a.resize(3);
            if (a[1].k != Value::STR) throw std::runtime_error("CheckNumberField: field name expected");
            return Values{Value::boolean(number_field(*o, a[1].s->c_str(), a[2]))};
        });

        // cameras.lua
        define(I, "Camera", [](Interp& I, Values& a) {
            auto s = self(a, "Camera");
            if (!string_field(I, *s, "type", Value::string("pinhole"), {"pinhole", "orthographic"}))
                exit_with(I, "Lua ERROR: Invalid camera type. Exiting...");
            if (!vec3_field(*s, "position", vec3_value({0, 0, 5}))) exit_with(I, "Lua ERROR: Invalid camera position vector. Exiting...");
            if (!vec3_field(*s, "look_at", vec3_value({0, 0, 0}))) exit_with(I, "Lua ERROR: Invalid camera 'look at' vector. Exiting...");
            if (!vec3_field(*s, "up", vec3_value({0, 1, 0}))) exit_with(I, "Lua ERROR: Invalid camera up vector. Exiting...");
            Value t = s->get("type");
            if (t.k == Value::STR && *t.s == "pinhole") {
                if (!number_field(*s, "fov", Value::integer(55))) exit_with(I, "Lua ERROR: Invalid camera FOV. Exiting...");
            } else {
                if (!number_field(*s, "min_x")) exit_with(I, "Lua ERROR: Invalid camera min_x. Exiting...");
                if (!number_field(*s, "max_x")) exit_with(I, "Lua ERROR: Invalid camera max_x. Exiting...");
                if (!number_field(*s, "min_y")) exit_with(I, "Lua ERROR: Invalid camera min_y. Exiting...");
                if (!number_field(*s, "max_y")) exit_with(I, "Lua ERROR: Invalid camera max_y. Exiting...");
            }
            Value v;
            v.k = Value::TABLE;
            v.t = s;
            reg(I).set("__CAMERA__", v);
            return Values{};
        });
        // pixel_samplers.lua
        define(I, "PixelSampler", [](Interp& I, Values& a) {
            auto s = self(a, "PixelSampler");
            if (!string_field(I, *s, "type", Value::string("jittered"), {"uniform", "regular", "jittered"}))
                exit_with(I, "Lua ERROR: Invalid pixel sampler type. Exiting...");
            if (!number_field(*s, "spp", Value::integer(16))) exit_with(I, "Lua ERROR: Invalid pixel sampler spp. Exiting...");
            Value v;
            v.k = Value::TABLE;
            v.t = s;
            reg(I).set("__PIXEL_SAMPLER__", v);
            return Values{};
        });
        // image_buffers.lua
        define(I, "ImageBuffer", [](Interp& I, Values& a) {
            auto s = self(a, "ImageBuffer");
            if (!number_field(*s, "width", Value::integer(512))) exit_with(I, "Lua ERROR: Invalid image buffer width. Exiting...");
            if (!number_field(*s, "height", Value::integer(512))) exit_with(I, "Lua ERROR: Invalid image buffer height. Exiting...");
            if (!string_field(I, *s, "file_name", Value::string("image.ppm")))
                exit_with(I, "Lua ERROR: Invalid image buffer file name. Exiting...");
            if (!number_field(*s, "viewport_top", Value::integer(0)))
                exit_with(I, "Lua ERROR: Invalid image buffer 'viewport_top'. Exiting...");
            if (!number_field(*s, "viewport_left", Value::integer(0)))
                exit_with(I, "Lua ERROR: Invalid image buffer 'viewport_left'. Exiting...");
            if (!number_field(*s, "viewport_width")) exit_with(I, "Lua ERROR: Invalid image buffer 'viewport_width'. Exiting...");
            if (!number_field(*s, "viewport_height")) exit_with(I, "Lua ERROR: Invalid image buffer 'viewport_height'. Exiting...");
            Value v;
            v.k = Value::TABLE;
            v.t = s;
            reg(I).set("__IMAGE_BUFFER__", v);
            return Values{};
        });
        // global_settings.lua
        define(I, "GlobalSettings", [](Interp& I, Values& a) {
            auto s = self(a, "GlobalSettings");
            if (!string_field(I, *s, "acceleration_structure", Value::string("bvh-sah"), {"none", "bvh-sah"}))
                exit_with(I, "Lua ERROR: Invalid acceleration structure. Exiting...");
            if (!vec3_field(*s, "background_color", vec3_value({0, 0, 0})))
                exit_with(I, "Lua ERROR: Invalid image background color. Exiting...");
            Value acc = s->get("acceleration_structure");
            if (acc.k == Value::STR && *acc.s == "bvh-sah")
                s->set("overlap_threshold", Value::number(std::numeric_limits<double>::infinity()));
            Value v;
            v.k = Value::TABLE;
            v.t = s;
            reg(I).set("__GLOBAL_SETTINGS__", v);
            return Values{};
        });
        // integrators.lua
        define(I, "Integrator", [](Interp& I, Values& a) {
            auto s = self(a, "Integrator");
            if (!string_field(I, *s, "type", Value::string("ray-caster"), {"ray-caster", "path-tracer"}))
                exit_with(I, "Lua ERROR: Invalid integrator type. Exiting...");
            auto is = [&](const char* f, const char* v) {
                Value x = s->get(f);
                return x.k == Value::STR && *x.s == v;
            };
            if (is("type", "path-tracer")) {
                if (!string_field(I, *s, "path_termination", Value::string("russian-roulette"), {"max-length", "russian-roulette"}))
                    exit_with(I, "Lua ERROR: Invalid integrator path termination criterion. Exiting...");
                if (!number_field(*s, "path_length", Value::integer(5)))
                    exit_with(I, "Lua ERROR: Invalid integrator path length. Exiting...");
            }
            if (is("type", "ray-caster")) {
                if (!string_field(I, *s, "pixel_value", Value::string("normal"), {"depth", "normal", "intersection-test-count"}))
                    exit_with(I, "Lua ERROR: Invalid ray caster type (pixel format). Exiting...");
            }
            if (is("type", "ray-caster")) {
                if (is("pixel_value", "depth")) {
                    if (s->get("minimum_depth").is_nil()) s->set("minimum_depth", Value::integer(-1));
                    if (s->get("maximum_depth").is_nil()) s->set("maximum_depth", Value::integer(-1));
                }
                if (is("pixel_value", "intersection-test-count")) {
                    if (s->get("minimum_count").is_nil()) s->set("minimum_count", Value::integer(-1));
                    if (s->get("maximum_count").is_nil()) s->set("maximum_count", Value::integer(-1));
                }
            }
            Value v;
            v.k = Value::TABLE;
            v.t = s;
            reg(I).set("__INTEGRATOR__", v);
            return Values{};
        });
        // bsdfs.lua
        Value samplers = Interp::make_table();
        samplers.t->set(Value::integer(1), Value::string("uniform"));
        samplers.t->set(Value::integer(2), Value::string("importance"));
        I.G->set("bsdf_samplers", samplers);
        auto bsdf = [](Interp&, Values& a, const char* fname, const char* type) {
            auto s = self(a, fname);
            s->set("type", Value::string(type));
            return s;
        };
        auto sampler_ok = [](Interp& I, Table& s) {
            return string_field(I, s, "bsdf_sampler", Value::string("importance"), {"uniform", "importance"});
        };
        define(I, "Lambertian", [bsdf, sampler_ok](Interp& I, Values& a) {
            auto s = bsdf(I, a, "Lambertian", "lambertian");
            if (!vec3_field(*s, "kd", vec3_value({0.75, 0.75, 0.75})))
                exit_with(I, "Lua ERROR: Invalid lambertian reflectance (kd field). Exiting...");
            if (!sampler_ok(I, *s)) exit_with(I, "Lua ERROR: Invalid lambertian BSDF sampler. Exiting...");
            return Values{Value::integer(add_to(I, "__BSDFS__", "__BSDF_COUNT__", a[0]))};
        });
        define(I, "SmoothConductor", [bsdf, sampler_ok](Interp& I, Values& a) {
            auto s = bsdf(I, a, "SmoothConductor", "smooth_conductor");
            if (!vec3_field(*s, "reflectance_at_normal_incidence", vec3_value({0.99, 0.99, 0.99})))
                exit_with(I, "Lua ERROR: Invalid smooth conductor reflectance at normal incidence. Exiting...");
            if (!sampler_ok(I, *s)) exit_with(I, "Lua ERROR: Invalid smooth conductor BSDF sampler. Exiting...");
            return Values{Value::integer(add_to(I, "__BSDFS__", "__BSDF_COUNT__", a[0]))};
        });
        define(I, "SmoothDielectric", [bsdf, sampler_ok](Interp& I, Values& a) {
            auto s = bsdf(I, a, "SmoothDielectric", "smooth_dielectric");
            if (!number_field(*s, "ior", Value::number(1.5))) exit_with(I, "Lua ERROR: Invalid smooth dielectric IOR. Exiting...");
            if (!sampler_ok(I, *s)) exit_with(I, "Lua ERROR: Invalid smooth dielectric BSDF sampler. Exiting...");
            return Values{Value::integer(add_to(I, "__BSDFS__", "__BSDF_COUNT__", a[0]))};
        });
        define(I, "CookTorrance", [bsdf, sampler_ok](Interp& I, Values& a) {
            auto s = bsdf(I, a, "CookTorrance", "cook_torrance");
            if (!number_field(*s, "roughness", Value::number(0.2)))
                exit_with(I, "Lua ERROR: Invalid Cook-Torrance roughness. Exiting...");
            if (!vec3_field(*s, "reflectance_at_normal_incidence", vec3_value({0.99, 0.99, 0.99})))
                exit_with(I, "Lua ERROR: Invalid Cook-Torrance reflectance at normal incidence. Exiting...");
            if (!sampler_ok(I, *s)) exit_with(I, "Lua ERROR: Invalid Cook-Torrance BSDF sampler. Exiting...");
            return Values{Value::integer(add_to(I, "__BSDFS__", "__BSDF_COUNT__", a[0]))};
        });
        // layered_bsdfs.lua
        define(I, "LayeredBSDF", [](Interp& I, Values& a) {
            auto s = self(a, "LayeredBSDF");
            const size_t n = s->length();
            s->set("num_layers", Value::integer((double)n));
            const Value cnt = reg(I).get("__BSDF_COUNT__");
            for (size_t i = 1; i <= n; ++i) {
                Value l = s->get(Value::integer((double)i));
                if (l.k != Value::NUM) exit_with(I, "Lua ERROR: Invalid layered BSDF layer. Exiting...");
                if (l.n < 1 || (cnt.k == Value::NUM && l.n > cnt.n))
                    exit_with(I, "Lua ERROR: Invalid layered BSDF layer: index out of range. Exiting...");
            }
            return Values{Value::integer(add_to(I, "__LAYERED_BSDFS__", "__LAYERED_BSDF_COUNT__", a[0]))};
        });
        // emissions.lua
        define(I, "Emission", [](Interp& I, Values& a) {
            auto s = self(a, "Emission");
            if (!vec3(a[0])) exit_with(I, "Lua ERROR: Invalid emission value. Exiting...");
            (void)s;
            return Values{Value::integer(add_to(I, "__EMISSIONS__", "__EMISSION_COUNT__", a[0]))};
        });
        // materials.lua
        define(I, "Material", [](Interp& I, Values& a) {
            auto s = self(a, "Material");
            if (!number_field(*s, "bsdf")) exit_with(I, "Lua ERROR: Invalid material layered BSDF. Exiting...");
            if (!number_field(*s, "emission")) exit_with(I, "Lua ERROR: Invalid material emission. Exiting...");
            Table& r = reg(I);
            const double b = s->get("bsdf").n, e = s->get("emission").n;
            // materials.lua compares the layered BSDF index against the BSDF count
            if (b < 1 || b > r.get("__BSDF_COUNT__").n)
                exit_with(I, "Lua ERROR: Invalid material layered BSDF: index out of range. Exiting...");
            if (e < 1 || e > r.get("__EMISSION_COUNT__").n)
                exit_with(I, "Lua ERROR: Invalid material emission: index out of range. Exiting...");
            return Values{Value::integer(add_to(I, "__MATERIALS__", "__MATERIAL_COUNT__", a[0]))};
        });
        // objects.lua
        auto object = [](Interp&, Values& a, const char* fname, const char* type) {
            auto s = self(a, fname);
            s->set("type", Value::string(type));
            return s;
        };
        define(I, "Triangle", [object](Interp& I, Values& a) {
            auto s = object(I, a, "Triangle", "triangle");
            if (!vec3_field(*s, "v1")) exit_with(I, "Lua ERROR: Invalid triangle vertex (v1). Exiting...");
            if (!vec3_field(*s, "v2")) exit_with(I, "Lua ERROR: Invalid triangle vertex (v2). Exiting...");
            if (!vec3_field(*s, "v3")) exit_with(I, "Lua ERROR: Invalid triangle vertex (v3). Exiting...");
            add_to(I, "__OBJECTS__", "__OBJECT_COUNT__", a[0]);
            return Values{};
        });
        define(I, "SmoothTriangle", [object](Interp& I, Values& a) {
            auto s = object(I, a, "SmoothTriangle", "smooth_triangle");
            if (!vec3_field(*s, "v1")) exit_with(I, "Lua ERROR: Invalid smooth triangle vertex (v1). Exiting...");
            if (!vec3_field(*s, "v2")) exit_with(I, "Lua ERROR: Invalid smooth triangle vertex (v2). Exiting...");
            if (!vec3_field(*s, "v3")) exit_with(I, "Lua ERROR: Invalid smooth triangle vertex (v3). Exiting...");
            if (!vec3_field(*s, "n1")) exit_with(I, "Lua ERROR: Invalid smooth triangle normal vector (n1). Exiting...");
            if (!vec3_field(*s, "n2")) exit_with(I, "Lua ERROR: Invalid smooth triangle normal vector (n2). Exiting...");
            if (!vec3_field(*s, "n3")) exit_with(I, "Lua ERROR: Invalid smooth triangle normal vector (n3). Exiting...");
            add_to(I, "__OBJECTS__", "__OBJECT_COUNT__", a[0]);
            return Values{};
        });
        define(I, "Sphere", [object](Interp& I, Values& a) {
            auto s = object(I, a, "Sphere", "sphere");
            if (!vec3_field(*s, "center")) exit_with(I, "Lua ERROR: Invalid sphere center. Exiting...");
            if (!number_field(*s, "radius", Value::integer(1))) exit_with(I, "Lua ERROR: Invalid sphere radius. Exiting...");
            add_to(I, "__OBJECTS__", "__OBJECT_COUNT__", a[0]);
            return Values{};
        });
        define(I, "Mesh", [object](Interp& I, Values& a) {
            auto s = object(I, a, "Mesh", "mesh");
            if (!string_field(I, *s, "filename")) exit_with(I, "Lua ERROR: Invalid mesh file name. Exiting...");
            add_to(I, "__OBJECTS__", "__OBJECT_COUNT__", a[0]);
            return Values{};
        });
    }
};

// Which lua_core modules the built-ins stand for. `require "core"` pulls in
// all of them (lua_core/core.lua); each registers its part.
bool builtin_module(Interp& I, const std::string& name) {
    static const std::unordered_set<std::string> parts = {
        "registervars", "validation", "materials", "bsdfs", "layered_bsdfs", "emissions", "objects",
        "cameras", "pixel_samplers", "image_buffers", "global_settings", "integrators"};
    if (name != "core" && !parts.count(name)) return false;
    if (name == "core" || name == "registervars") {
        if (I.loaded->get("registervars").is_nil()) {
            Core::registervars(I);
            I.loaded->set("registervars", Value::boolean(true));
        }
    }
    if (I.G->get("Camera").is_nil()) Core::install(I);
    if (name == "core")
        for (const std::string& p : parts) I.loaded->set(p, Value::boolean(true));
    return true;
}

// ===========================================================================
// Registry dump (JSON, for inspection and tests)
// ===========================================================================

void to_json(const Interp& I, const Value& v, std::string& out, std::unordered_set<const Table*>& seen) {
    switch (v.k) {
        case Value::NIL: out += "null"; return;
        case Value::BOOL: out += v.b ? "true" : "false"; return;
        case Value::NUM: {
            if (std::isinf(v.n)) {
                out += v.n > 0 ? "\"inf\"" : "\"-inf\"";
                return;
            }
            char buf[64];
            if (v.isint) {
                std::snprintf(buf, sizeof buf, "%lld", (long long)v.n);
            } else {
                std::snprintf(buf, sizeof buf, "%.17g", v.n);
                if (!std::strpbrk(buf, ".eEn")) std::strcat(buf, ".0");
            }
            out += buf;
            return;
        }
        case Value::STR: {
            out += '"';
            for (char c : *v.s) {
                if (c == '"' || c == '\\') out += '\\', out += c;
                else if ((unsigned char)c < 0x20) {
                    char buf[8];
                    std::snprintf(buf, sizeof buf, "\\u%04x", (unsigned char)c);
                    out += buf;
                } else out += c;
            }
            out += '"';
            return;
        }
        case Value::FUNC: out += "\"<function>\""; return;
        case Value::TABLE: {
            if (seen.count(v.t.get())) {
                out += "\"<cycle>\"";
                return;
            }
            seen.insert(v.t.get());
            // keys sorted (array part by index, then strings) so dumps compare
            std::map<std::string, Value> items;
            Value k, val;
            Value cur;
            while (v.t->next(cur, k, val)) {
                std::string key = k.k == Value::STR ? *k.s : I.tostring(k);
                if (k.k == Value::NUM) {
                    char buf[32];
                    std::snprintf(buf, sizeof buf, "#%012lld", (long long)k.n);
                    key = buf;
                }
                items.emplace(key, val);
                cur = k;
            }
            out += '{';
            bool first = true;
            for (auto& kv : items) {
                if (!first) out += ',';
                first = false;
                to_json(I, Value::string(kv.first), out, seen);
                out += ':';
                to_json(I, kv.second, out, seen);
            }
            out += '}';
            seen.erase(v.t.get());
            return;
        }
    }
}

// ===========================================================================
// Extraction (lua_bind/lua_bind.cpp) and scene assembly
// (core/swpathtracer_app.cpp)
// ===========================================================================

struct Reader {
    const Table& t;
    // getNumberField: only a Lua number counts (lua_bind.h:57-67)
    bool num(const char* f, double& out) const {
        Value v = t.get(f);
        if (v.k != Value::NUM) return false;
        out = v.n;
        return true;
    }
    bool str(const char* f, std::string& out) const {
        Value v = t.get(f);
        if (v.k != Value::STR) return false;
        out = *v.s;
        return true;
    }
    // getVec3Field: any table; entries that are not numbers read as 0
    // (lua_tonumber), lua_bind.cpp:511-526
    bool vec3(const char* f, float out[3]) const {
        Value v = t.get(f);
        if (v.k != Value::TABLE) return false;
        for (int i = 0; i < 3; ++i) {
            Value e = v.t->get(Value::number(i + 1));
            out[i] = e.k == Value::NUM ? (float)e.n : 0.0f;
        }
        return true;
    }
};

const Table& element(const Table& reg, const char* name, const char* what) {
    Value v = reg.get(name);
    if (v.k != Value::TABLE) throw std::runtime_error(std::string("ERROR: Invalid ") + what + " definition. Exiting...");
    return *v.t;
}
// lua_tonumber of a registry count (0 when absent)
size_t count_of(const Table& reg, const char* name) {
    Value v = reg.get(name);
    return v.k == Value::NUM && v.n > 0 ? (size_t)v.n : 0;
}
const Table& item(const Table& reg, const char* list, size_t i, const char* what) {
    Value l = reg.get(list);
    Value v = l.k == Value::TABLE ? l.t->get(Value::number((double)i)) : Value();
    if (v.k != Value::TABLE) throw std::runtime_error(std::string("ERROR: Invalid ") + what + ". Exiting...");
    return *v.t;
}
// lua_bind stores numbers into unsigned / size_t fields: a C conversion of
// the double (negative values wrap through the integer type)
uint32_t to_u32(double d) { return (uint32_t)(int64_t)d; }

void extract(const Interp& I, const std::string& workdir, SceneOut& out) {
    Value rv = I.G->get("__RENDERER_ELEMENTS__");
    if (rv.k != Value::TABLE) throw std::runtime_error("ERROR: the script defined no scene (no __RENDERER_ELEMENTS__)");
    const Table& reg = *rv.t;
    spt_lua_info& info = out.info;

    // getImageBuffer (lua_bind.cpp:33-57)
    {
        Reader r{element(reg, "__IMAGE_BUFFER__", "image buffer")};
        double w, h, vt, vl, vw, vh;
        std::string fname;
        if (!r.num("width", w) || !r.num("height", h) || !r.num("viewport_top", vt) || !r.num("viewport_left", vl) ||
            !r.num("viewport_width", vw) || !r.num("viewport_height", vh) || !r.str("file_name", fname))
            throw std::runtime_error("ERROR: Invalid image buffer definition. Exiting...");
        info.camera.width = to_u32(w);
        info.camera.height = to_u32(h);
        info.params.viewport[0] = to_u32(vt);
        info.params.viewport[1] = to_u32(vl);
        info.params.viewport[2] = to_u32(vw);
        info.params.viewport[3] = to_u32(vh);
        std::snprintf(info.file_name, sizeof info.file_name, "%s", fname.c_str());
    }
    // getCamera (lua_bind.cpp:59-95)
    {
        Reader r{element(reg, "__CAMERA__", "camera")};
        std::string type;
        if (!r.vec3("position", info.camera.position) || !r.vec3("look_at", info.camera.look_at) ||
            !r.vec3("up", info.camera.up) || !r.str("type", type))
            throw std::runtime_error("ERROR: Invalid camera definition. Exiting...");
        if (type == "orthographic") {
            double a = -1, b = 1, c = -1, d = 1;
            if (!r.num("min_x", a) || !r.num("max_x", b) || !r.num("min_y", c) || !r.num("max_y", d))
                throw std::runtime_error("ERROR: Invalid camera definition. Exiting...");
            info.camera.type = SPT_CAMERA_ORTHOGRAPHIC;
            info.camera.min_x = (float)a;
            info.camera.max_x = (float)b;
            info.camera.min_y = (float)c;
            info.camera.max_y = (float)d;
        } else if (type == "pinhole") {
            double fov = 55;
            if (!r.num("fov", fov)) throw std::runtime_error("ERROR: Invalid camera definition. Exiting...");
            info.camera.type = SPT_CAMERA_PINHOLE;
            info.camera.fov_deg = (float)fov;
        } else {
            throw std::runtime_error("ERROR: Invalid camera definition. Exiting...");  // the reference leaves it null
        }
    }
    // getGlobalSettings (lua_bind.cpp:122-155)
    {
        Reader r{element(reg, "__GLOBAL_SETTINGS__", "global settings")};
        std::string acc;
        if (!r.vec3("background_color", info.params.background) || !r.str("acceleration_structure", acc))
            throw std::runtime_error("ERROR: Invalid global settings definition. Exiting...");
        if (acc == "bvh-sah") {
            info.accel = SPT_ACCEL_BVH_SAH;
            info.overlap_threshold = std::numeric_limits<float>::infinity();
        } else if (acc == "sbvh-sah") {
            double t = 0;
            if (!r.num("overlap_threshold", t))
                throw std::runtime_error("ERROR: Invalid global settings definition. Exiting...");
            info.accel = SPT_ACCEL_SBVH_SAH;
            info.overlap_threshold = (float)t;
        } else {
            info.accel = SPT_ACCEL_NONE;
        }
    }
    // getIntegrator (lua_bind.cpp:458-508) + loadIntegrator (swpathtracer_app.cpp:221-259)
    {
        Reader r{element(reg, "__INTEGRATOR__", "integrator")};
        std::string type;
        if (!r.str("type", type)) throw std::runtime_error("ERROR: Invalid itegrator definiton. Exiting...");
        info.min_value = info.max_value = 0.0f;
        info.params.max_depth = -1;
        info.params.rr_depth = -1;
        if (type == "ray-caster") {
            std::string kind;
            if (!r.str("pixel_value", kind)) throw std::runtime_error("ERROR: Invalid itegrator definiton. Exiting...");
            if (kind == "normal") {
                info.integrator = SPT_LUA_RAYCAST_NORMAL;
            } else if (kind == "depth") {
                info.integrator = SPT_LUA_RAYCAST_DEPTH;
                double lo, hi;
                if (!r.num("minimum_depth", lo) || !r.num("maximum_depth", hi))
                    throw std::runtime_error("ERROR: Invalid itegrator definiton. Exiting...");
                info.min_value = (float)lo;
                info.max_value = (float)hi;
            } else if (kind == "intersection-test-count") {
                info.integrator = SPT_LUA_RAYCAST_COUNT;
                double lo, hi;
                if (!r.num("minimum_count", lo) || !r.num("maximum_count", hi))
                    throw std::runtime_error("ERROR: Invalid itegrator definiton. Exiting...");
                info.min_value = (float)lo;
                info.max_value = (float)hi;
            } else {
                info.integrator = SPT_LUA_NONE;
            }
        } else if (type == "path-tracer") {
            info.integrator = SPT_LUA_PATHTRACER;
            std::string term;
            if (!r.str("path_termination", term)) throw std::runtime_error("ERROR: Invalid itegrator definiton. Exiting...");
            double len;
            if (!r.num("path_length", len)) throw std::runtime_error("ERROR: Invalid integrator path length. Exiting...");
            const int32_t pl = (int32_t)(int64_t)len;
            info.path_length = pl;
            if (term == "max-length") {
                info.params.max_depth = pl;  // pathtracer.cpp:18
            } else {
                info.params.rr_depth = pl;   // pathtracer.cpp:43
            }
            // getPixelSampler (lua_bind.cpp:97-120), read for path tracers only
            Reader ps{element(reg, "__PIXEL_SAMPLER__", "pixel sampler")};
            std::string pst;
            double spp = 4;
            if (!ps.str("type", pst) || !ps.num("spp", spp))
                throw std::runtime_error("ERROR: Invalid pixel sampler definition. Exiting...");
            if (pst == "uniform") info.params.pixel_sampler = SPT_PIXEL_UNIFORM;
            else if (pst == "regular") info.params.pixel_sampler = SPT_PIXEL_REGULAR;
            else if (pst == "jittered") info.params.pixel_sampler = SPT_PIXEL_JITTERED;
            else throw std::runtime_error("ERROR: Invalid pixel sampler definition. Exiting...");
            info.params.spp = (uint32_t)(uint64_t)spp;
        } else {
            info.integrator = SPT_LUA_NONE;
        }
        info.params.flags = SPT_RENDER_FILM32;
        info.params.rng = SPT_RNG_COUNTER;
    }

    // getBSDFAt (lua_bind.cpp:171-226) + loadBSDFs (swpathtracer_app.cpp:88-147)
    const size_t n_bsdf = count_of(reg, "__BSDF_COUNT__");
    if (!n_bsdf) throw std::runtime_error("ERROR: No BSDF defined. Exiting...");
    std::vector<spt_layer> bsdfs(n_bsdf);
    for (size_t i = 0; i < n_bsdf; ++i) {
        Reader r{item(reg, "__BSDFS__", i + 1, "BSDF definition")};
        std::string type, sampler;
        if (!r.str("type", type) || !r.str("bsdf_sampler", sampler))
            throw std::runtime_error("ERROR: Invalid BSDF definition. Exiting...");
        spt_layer& L = bsdfs[i];
        std::memset(&L, 0, sizeof L);
        // only "uniform" selects UniformBSDFSampler; anything else the BSDF's own
        L.sampler = sampler == "uniform" ? SPT_SAMPLER_UNIFORM : SPT_SAMPLER_IMPORTANCE;
        if (type == "lambertian") {
            L.type = SPT_BSDF_LAMBERTIAN;
            if (!r.vec3("kd", L.kd)) throw std::runtime_error("ERROR: Invalid BSDF definition. Exiting...");
        } else if (type == "smooth_conductor") {
            L.type = SPT_BSDF_SMOOTH_CONDUCTOR;
            if (!r.vec3("reflectance_at_normal_incidence", L.r0))
                throw std::runtime_error("ERROR: Invalid BSDF definition. Exiting...");
        } else if (type == "smooth_dielectric") {
            L.type = SPT_BSDF_SMOOTH_DIELECTRIC;
            L.sampler = SPT_SAMPLER_IMPORTANCE;  // always its reflection/refraction samplers (app:119-127)
            double ior = 1.49;
            if (!r.num("ior", ior)) throw std::runtime_error("ERROR: Invalid BSDF definition. Exiting...");
            L.ior = (float)ior;
        } else if (type == "cook_torrance") {
            L.type = SPT_BSDF_COOK_TORRANCE;
            double m = 0.2;
            if (!r.num("roughness", m) || !r.vec3("reflectance_at_normal_incidence", L.r0))
                throw std::runtime_error("ERROR: Invalid BSDF definition. Exiting...");
            L.roughness = (float)m;
        } else {
            throw std::runtime_error("ERROR: Invalid BSDF definition (unknown type '" + type + "'). Exiting...");
        }
    }
    // getLayeredBSDFAt (lua_bind.cpp:242-266)
    const size_t n_layered = count_of(reg, "__LAYERED_BSDF_COUNT__");
    if (!n_layered) throw std::runtime_error("ERROR: No layered BSDF defined. Exiting...");
    std::vector<std::vector<uint32_t>> layered(n_layered);
    for (size_t i = 0; i < n_layered; ++i) {
        const Table& t = item(reg, "__LAYERED_BSDFS__", i + 1, "layered BSDF definition");
        double nl;
        if (!Reader{t}.num("num_layers", nl)) throw std::runtime_error("ERROR: Invalid layered BSDF definition. Exiting...");
        for (uint32_t k = 1; k <= to_u32(nl); ++k) {
            Value v = t.get(Value::number(k));
            const int64_t id = v.k == Value::NUM ? (int64_t)v.n : 0;
            if (id < 1 || (size_t)id > n_bsdf)
                throw std::runtime_error("ERROR: Invalid layered BSDF definition (layer index). Exiting...");
            layered[i].push_back((uint32_t)(id - 1));
        }
    }
    // getEmissionAt (lua_bind.cpp:282-302)
    const size_t n_em = count_of(reg, "__EMISSION_COUNT__");
    std::vector<std::array<float, 3>> emissions(n_em);
    for (size_t i = 0; i < n_em; ++i) {
        const Table& t = item(reg, "__EMISSIONS__", i + 1, "emission");
        for (int c = 0; c < 3; ++c) {
            Value v = t.get(Value::number(c + 1));
            emissions[i][c] = v.k == Value::NUM ? (float)v.n : 0.0f;
        }
    }
    // getMaterialAt (lua_bind.cpp:318-340) + loadMaterials (swpathtracer_app.cpp:171-179)
    const size_t n_mat = count_of(reg, "__MATERIAL_COUNT__");
    if (!n_mat) throw std::runtime_error("ERROR: No material defined. Exiting...");
    out.mats.assign(n_mat, spt_material{});
    for (size_t i = 0; i < n_mat; ++i) {
        Reader r{item(reg, "__MATERIALS__", i + 1, "material")};
        double b, e;
        if (!r.num("bsdf", b) || !r.num("emission", e)) throw std::runtime_error("ERROR: Invalid material. Exiting...");
        const int64_t bi = (int64_t)b - 1, ei = (int64_t)e - 1;
        if (bi < 0 || (size_t)bi >= n_layered || ei < 0 || (size_t)ei >= n_em)
            throw std::runtime_error("ERROR: Invalid material (index out of range). Exiting...");
        spt_material& M = out.mats[i];
        std::memset(&M, 0, sizeof M);
        const auto& ids = layered[(size_t)bi];
        if (ids.size() > SPT_MAX_LAYERS)
            throw std::runtime_error("ERROR: a layered BSDF has more than " + std::to_string(SPT_MAX_LAYERS) +
                                     " layers (the device material record's limit)");
        M.n_layers = (uint32_t)ids.size();
        for (size_t k = 0; k < ids.size(); ++k) M.layers[k] = bsdfs[ids[k]];
        for (int c = 0; c < 3; ++c) M.emission[c] = emissions[(size_t)ei][c];
    }
    // getObjectAt (lua_bind.cpp:356-456) + loadPrimitives (swpathtracer_app.cpp:181-218)
    const size_t n_obj = count_of(reg, "__OBJECT_COUNT__");
    if (!n_obj) throw std::runtime_error("ERROR: No object (primitive) defined. Exiting...");
    auto push = [&](uint8_t kind, const float p[9], const float* nrm, uint32_t mat) {
        out.prims.insert(out.prims.end(), p, p + 9);
        out.kinds.push_back(kind);
        for (int k = 0; k < 9; ++k) out.normals.push_back(nrm ? nrm[k] : 0.0f);
        out.mat_ids.push_back(mat);
    };
    for (size_t i = 0; i < n_obj; ++i) {
        const Table& t = item(reg, "__OBJECTS__", i + 1, "object definiton");
        Reader r{t};
        double mat;
        if (!r.num("material", mat)) throw std::runtime_error("ERROR: Invalid object definiton. Exiting...");
        const int64_t mi = (int64_t)mat - 1;
        if (mi < 0 || (size_t)mi >= n_mat)
            throw std::runtime_error("ERROR: Invalid object definiton (material index out of range). Exiting...");
        std::string type;
        if (!r.str("type", type)) throw std::runtime_error("ERROR: Invalid object definition. Exiting...");
        float p[9] = {0}, n[9] = {0};
        if (type == "triangle") {
            if (!r.vec3("v1", p) || !r.vec3("v2", p + 3) || !r.vec3("v3", p + 6))
                throw std::runtime_error("ERROR: Invalid object definition. Exiting...");
            push(SPT_PRIM_TRIANGLE, p, nullptr, (uint32_t)mi);
        } else if (type == "smooth_triangle") {
            if (!r.vec3("v1", p) || !r.vec3("v2", p + 3) || !r.vec3("v3", p + 6) || !r.vec3("n1", n) ||
                !r.vec3("n2", n + 3) || !r.vec3("n3", n + 6))
                throw std::runtime_error("ERROR: Invalid object definition. Exiting...");
            push(SPT_PRIM_TRIANGLE, p, n, (uint32_t)mi);
        } else if (type == "sphere") {
            double rad = 1.0;
            if (!r.vec3("center", p) || !r.num("radius", rad))
                throw std::runtime_error("ERROR: Invalid object definiton. Exiting...");
            p[3] = (float)rad;
            push(SPT_PRIM_SPHERE, p, nullptr, (uint32_t)mi);
        } else if (type == "mesh") {
            std::string file, sub;
            if (!r.str("filename", file) || !r.str("render_submeshes", sub))
                throw std::runtime_error("ERROR: Invalid object definiton. Exiting...");
            const bool all = sub == "all";
            std::unordered_map<std::string, uint32_t> submeshes;
            Value sm = t.get("sub_meshes");
            if (sm.k == Value::TABLE) {
                Value k, v, cur;
                while (sm.t->next(cur, k, v)) {
                    cur = k;
                    if (v.k != Value::TABLE) throw std::runtime_error("ERROR: Invalid object definiton. Exiting...");
                    Reader s{*v.t};
                    std::string name;
                    double m2;
                    if (!s.str("name", name) || !s.num("material", m2))
                        throw std::runtime_error("ERROR: Invalid object definiton. Exiting...");
                    const int64_t m2i = (int64_t)m2 - 1;
                    if (m2i < 0 || (size_t)m2i >= n_mat)
                        throw std::runtime_error("ERROR: Invalid object definiton (sub-mesh material). Exiting...");
                    submeshes.emplace(name, (uint32_t)m2i);  // first insertion wins, as unordered_map::insert
                }
            }
            // Scene::loadMesh (core/scene.cpp:32-68)
            for (const ObjMesh& m : read_obj(join_path(workdir, file))) {
                auto it = submeshes.find(m.name);
                if (it == submeshes.end() && !all) continue;
                const uint32_t mm = it == submeshes.end() ? (uint32_t)mi : it->second;
                for (size_t f = 0; f * 9 < m.tris.size(); ++f) push(SPT_PRIM_TRIANGLE, &m.tris[9 * f], &m.normals[9 * f], mm);
            }
        } else {
            throw std::runtime_error("ERROR: Invalid object definition (unknown type '" + type + "'). Exiting...");
        }
    }
    info.n_prims = out.kinds.size();
    info.n_materials = out.mats.size();
    info.n_objects = n_obj;
    info.n_bsdfs = n_bsdf;
}

}  // namespace

// ===========================================================================
// OBJ reader: Assimp's view of an OBJ file as Scene::loadMesh consumes it
// (aiProcess_Triangulate: fans; every face corner its own vertex; meshes in
// file order, a new one at each `o` / `g` followed by faces; faces whose
// corners repeat a position index are dropped, as aiProcess_FindDegenerates
// turns them into points or lines).
// ===========================================================================

std::vector<ObjMesh> read_obj(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Unable to open file \"" + path + "\".");
    std::vector<std::array<float, 3>> pos, nrm;
    std::vector<ObjMesh> meshes;
    std::string name = "default";
    bool open = false;
    auto parse3 = [&](std::istringstream& ss, std::array<float, 3>& v) {
        for (int i = 0; i < 3; ++i) {
            std::string tok;
            if (!(ss >> tok)) throw std::runtime_error("OBJ: malformed vertex line in " + path);
            v[i] = std::strtof(tok.c_str(), nullptr);
        }
    };
    auto idx = [](const std::string& tok, size_t n) -> int64_t {
        const long long i = std::strtoll(tok.c_str(), nullptr, 10);
        return i > 0 ? i - 1 : (int64_t)n + i;
    };
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        std::string cmd;
        if (!(ss >> cmd) || cmd[0] == '#') continue;
        if (cmd == "v") {
            std::array<float, 3> v;
            parse3(ss, v);
            pos.push_back(v);
        } else if (cmd == "vn") {
            std::array<float, 3> v;
            parse3(ss, v);
            nrm.push_back(v);
        } else if (cmd == "o" || cmd == "g") {
            std::string rest, w;
            while (ss >> w) rest += (rest.empty() ? "" : " ") + w;
            name = rest.empty() ? "default" : rest;
            open = false;
        } else if (cmd == "f") {
            if (!open) {
                meshes.push_back(ObjMesh{name, {}, {}});
                open = true;
            }
            std::vector<std::pair<int64_t, int64_t>> corners;
            std::string c;
            while (ss >> c) {
                const size_t s1 = c.find('/');
                const int64_t vi = idx(c.substr(0, s1), pos.size());
                int64_t ni = -1;
                if (s1 != std::string::npos) {
                    const size_t s2 = c.find('/', s1 + 1);
                    if (s2 != std::string::npos && s2 + 1 < c.size()) ni = idx(c.substr(s2 + 1), nrm.size());
                }
                if (vi < 0 || (size_t)vi >= pos.size()) throw std::runtime_error("OBJ: vertex index out of range in " + path);
                if (s1 != std::string::npos && ni >= 0 && (size_t)ni >= nrm.size())
                    throw std::runtime_error("OBJ: normal index out of range in " + path);
                corners.emplace_back(vi, ni);
            }
            ObjMesh& m = meshes.back();
            for (size_t k = 1; k + 1 < corners.size(); ++k) {
                const std::pair<int64_t, int64_t> tri[3] = {corners[0], corners[k], corners[k + 1]};
                if (tri[0].first == tri[1].first || tri[1].first == tri[2].first || tri[0].first == tri[2].first)
                    continue;
                const bool has_n = tri[0].second >= 0 && tri[1].second >= 0 && tri[2].second >= 0;
                for (int j = 0; j < 3; ++j)
                    for (int a = 0; a < 3; ++a) {
                        m.tris.push_back(pos[(size_t)tri[j].first][a]);
                        m.normals.push_back(has_n ? nrm[(size_t)tri[j].second][a] : 0.0f);
                    }
            }
        }
    }
    std::vector<ObjMesh> out;
    for (ObjMesh& m : meshes)
        if (!m.tris.empty()) out.push_back(std::move(m));
    return out;
}

void load_scene(const std::string& script, const std::string& workdir, uint32_t flags, SceneOut& out) {
    Interp I;
    I.G = std::make_shared<Table>();
    I.loaded = std::make_shared<Table>();
    I.workdir = workdir;
    open_libs(I);
    I.builtin_module = builtin_module;
    I.builtin_only = (flags & SPT_LUA_BUILTIN_CORE) != 0;
    struct Cleanup {
        Interp& I;
        ~Cleanup() {  // break the _G <-> _G._G cycle and closure/env cycles
            I.G->entries.clear();
            I.G->index.clear();
            I.G->arr.clear();
            I.loaded->entries.clear();
            I.loaded->index.clear();
        }
    } cleanup{I};
    try {
        std::string chunkname = script;
        const size_t sl = chunkname.find_last_of('/');
        if (sl != std::string::npos) chunkname = chunkname.substr(sl + 1);
        I.run_file(join_path(workdir, script), chunkname, {});
    } catch (const ExitRequest&) {
        out.printed = I.printed;
        throw std::runtime_error(I.last_line.empty() ? std::string("os.exit() called by the script") : I.last_line);
    }
    out.printed = I.printed;
    std::unordered_set<const Table*> seen;
    out.registry_json.clear();
    to_json(I, I.G->get("__RENDERER_ELEMENTS__"), out.registry_json, seen);
    extract(I, workdir, out);
}

}  // namespace spt_lua

// ===========================================================================
// C ABI
// ===========================================================================

struct spt_lua_scene {
    spt_lua::SceneOut s;
};

namespace {
spt_status copy_text(const std::string& text, char* buf, uint64_t* len) {
    if (!len) return SPT_ERR_INVALID;
    const uint64_t need = (uint64_t)text.size() + 1;
    if (buf && *len >= need) std::memcpy(buf, text.c_str(), need);
    const bool fits = buf == nullptr || *len >= need;
    *len = need;
    return fits ? SPT_OK : SPT_ERR_OVERFLOW;
}
}  // namespace

extern "C" {

spt_status spt_lua_load(const char* script, const char* workdir, uint32_t flags, spt_lua_scene** out, char* err,
                        size_t err_len) {
    if (err && err_len) err[0] = '\0';
    if (!script || !out) return SPT_ERR_INVALID;
    *out = nullptr;
    auto sc = std::make_unique<spt_lua_scene>();
    try {
        spt_lua::load_scene(script, workdir ? workdir : "", flags, sc->s);
    } catch (const std::exception& e) {
        if (err && err_len) std::snprintf(err, err_len, "%s", e.what());
        return SPT_ERR_INVALID;
    }
    *out = sc.release();
    return SPT_OK;
}

spt_status spt_lua_get_info(const spt_lua_scene* sc, spt_lua_info* info) {
    if (!sc || !info) return SPT_ERR_INVALID;
    *info = sc->s.info;
    return SPT_OK;
}

spt_status spt_lua_get_prims(const spt_lua_scene* sc, float* prims, uint8_t* kinds, float* normals,
                             uint32_t* mat_ids) {
    if (!sc) return SPT_ERR_INVALID;
    const spt_lua::SceneOut& s = sc->s;
    if (prims && !s.prims.empty()) std::memcpy(prims, s.prims.data(), s.prims.size() * sizeof(float));
    if (kinds && !s.kinds.empty()) std::memcpy(kinds, s.kinds.data(), s.kinds.size());
    if (normals && !s.normals.empty()) std::memcpy(normals, s.normals.data(), s.normals.size() * sizeof(float));
    if (mat_ids && !s.mat_ids.empty()) std::memcpy(mat_ids, s.mat_ids.data(), s.mat_ids.size() * sizeof(uint32_t));
    return SPT_OK;
}

spt_status spt_lua_get_materials(const spt_lua_scene* sc, spt_material* mats) {
    if (!sc || !mats) return SPT_ERR_INVALID;
    if (!sc->s.mats.empty()) std::memcpy(mats, sc->s.mats.data(), sc->s.mats.size() * sizeof(spt_material));
    return SPT_OK;
}

spt_status spt_lua_registry_json(const spt_lua_scene* sc, char* buf, uint64_t* len) {
    if (!sc) return SPT_ERR_INVALID;
    return copy_text(sc->s.registry_json, buf, len);
}

spt_status spt_lua_printed(const spt_lua_scene* sc, char* buf, uint64_t* len) {
    if (!sc) return SPT_ERR_INVALID;
    return copy_text(sc->s.printed, buf, len);
}

void spt_lua_destroy(spt_lua_scene* sc) { delete sc; }

}  // extern "C"
