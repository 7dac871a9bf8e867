// Falm C++ host API: a small JSON value for the setup, turbine, index and
// AP files (the reference parses them with nlohmann::json, vendored at
// src/nlohmann/json.hpp; src/typedef.h names it Json).
//
// Covers what the reference's setup handling uses (src/falm.h:222-340,
// src/alm/alm.h:20-57, src/alm/turbineHandler.h:121-205,
// src/alm/bladeHandler.h:16-104, src/FalmIO.h:19-67): parse a document,
// read members with operator[] / contains / size / is_* / get<T>() or an
// implicit conversion, build objects and arrays, and dump them.  Json keeps
// object keys sorted on dump (as nlohmann::json does); OrderedJson keeps
// insertion order (nlohmann::ordered_json, used for the AP file).
#ifndef FALM_AMD_JSON_H
#define FALM_AMD_JSON_H

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <istream>
#include <iterator>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

namespace Falm {

template <bool ORDERED>
class BasicJson {
public:
    enum class Kind { Null, Bool, Number, String, Array, Object };

    BasicJson() {}
    BasicJson(std::nullptr_t) {}
    BasicJson(bool b) : kind_(Kind::Bool), b_(b) {}
    template <class T, typename std::enable_if<std::is_arithmetic<T>::value && !std::is_same<T, bool>::value,
                                               int>::type = 0>
    BasicJson(T v) : kind_(Kind::Number), num_((double)v), integral_(std::is_integral<T>::value) {}
    BasicJson(const char *s) : kind_(Kind::String), str_(s) {}
    BasicJson(const std::string &s) : kind_(Kind::String), str_(s) {}
    // {a, b, c}: an array (an object is built member by member)
    BasicJson(std::initializer_list<BasicJson> l) : kind_(Kind::Array), arr_(l) {}
    template <class T>
    BasicJson(const std::vector<T> &v) : kind_(Kind::Array) {
        for (const T &x : v) arr_.push_back(BasicJson(x));
    }

    static BasicJson array() {
        BasicJson j;
        j.kind_ = Kind::Array;
        return j;
    }
    static BasicJson object() {
        BasicJson j;
        j.kind_ = Kind::Object;
        return j;
    }

    // ---------------------------------------------------------- parsing
    static BasicJson parse(const std::string &text) {
        Parser p{text.c_str(), text.c_str() + text.size()};
        p.ws();
        BasicJson j = p.value();
        p.ws();
        if (p.s != p.e) p.fail("trailing characters");
        return j;
    }
    static BasicJson parse(std::istream &in) {
        std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
        return parse(text);
    }

    // ----------------------------------------------------------- access
    Kind kind() const { return kind_; }
    bool is_null() const { return kind_ == Kind::Null; }
    bool is_boolean() const { return kind_ == Kind::Bool; }
    bool is_number() const { return kind_ == Kind::Number; }
    bool is_number_integer() const { return kind_ == Kind::Number && integral_; }
    bool is_string() const { return kind_ == Kind::String; }
    bool is_array() const { return kind_ == Kind::Array; }
    bool is_object() const { return kind_ == Kind::Object; }

    size_t size() const {
        if (kind_ == Kind::Array) return arr_.size();
        if (kind_ == Kind::Object) return obj_.size();
        return kind_ == Kind::Null ? 0 : 1;
    }
    bool contains(const std::string &key) const { return kind_ == Kind::Object && find(key) != nullptr; }

    const BasicJson &operator[](const std::string &key) const {
        if (kind_ != Kind::Object) throw std::runtime_error("json: member access on a non-object (" + key + ")");
        const BasicJson *v = find(key);
        if (!v) throw std::runtime_error("json: no member '" + key + "'");
        return *v;
    }
    const BasicJson &operator[](const char *key) const { return (*this)[std::string(key)]; }
    BasicJson &operator[](const std::string &key) {
        if (kind_ == Kind::Null) kind_ = Kind::Object;
        if (kind_ != Kind::Object) throw std::runtime_error("json: member access on a non-object (" + key + ")");
        for (auto &kv : obj_)
            if (kv.first == key) return kv.second;
        obj_.emplace_back(key, BasicJson());
        return obj_.back().second;
    }
    BasicJson &operator[](const char *key) { return (*this)[std::string(key)]; }
    template <class I, typename std::enable_if<std::is_integral<I>::value, int>::type = 0>
    const BasicJson &operator[](I i) const {
        if (kind_ != Kind::Array || (size_t)i >= arr_.size()) throw std::runtime_error("json: index out of range");
        return arr_[(size_t)i];
    }
    template <class I, typename std::enable_if<std::is_integral<I>::value, int>::type = 0>
    BasicJson &operator[](I i) {
        if (kind_ != Kind::Array || (size_t)i >= arr_.size()) throw std::runtime_error("json: index out of range");
        return arr_[(size_t)i];
    }
    void push_back(const BasicJson &v) {
        if (kind_ == Kind::Null) kind_ = Kind::Array;
        if (kind_ != Kind::Array) throw std::runtime_error("json: push_back on a non-array");
        arr_.push_back(v);
    }
    const std::vector<BasicJson> &items() const { return arr_; }
    const std::vector<std::pair<std::string, BasicJson>> &members() const { return obj_; }

    template <class T>
    T get() const {
        return Get<T>::from(*this);
    }
    // implicit conversion to numbers, bool and std::string (not to char,
    // so that assigning to a std::string is unambiguous)
    template <class T, typename std::enable_if<(std::is_arithmetic<T>::value && !std::is_same<T, char>::value) ||
                                                   std::is_same<T, std::string>::value,
                                               int>::type = 0>
    operator T() const {
        return get<T>();
    }

    bool operator==(const BasicJson &o) const {
        if (kind_ != o.kind_) return false;
        switch (kind_) {
        case Kind::Null: return true;
        case Kind::Bool: return b_ == o.b_;
        case Kind::Number: return num_ == o.num_;
        case Kind::String: return str_ == o.str_;
        case Kind::Array: return arr_ == o.arr_;
        case Kind::Object: {
            if (obj_.size() != o.obj_.size()) return false;
            for (const auto &kv : obj_) {
                const BasicJson *v = o.find(kv.first);
                if (!v || !(*v == kv.second)) return false;
            }
            return true;
        }
        }
        return false;
    }
    bool operator!=(const BasicJson &o) const { return !(*this == o); }
    template <class T, typename std::enable_if<std::is_arithmetic<T>::value && !std::is_same<T, bool>::value,
                                               int>::type = 0>
    bool operator!=(T v) const {
        return !(is_number() && num_ == (double)v);
    }

    // ----------------------------------------------------------- output
    std::string dump(int indent = -1) const {
        std::string out;
        write(out, indent, 0);
        return out;
    }

private:
    Kind kind_ = Kind::Null;
    bool b_ = false;
    double num_ = 0.0;
    bool integral_ = false;
    std::string str_;
    std::vector<BasicJson> arr_;
    std::vector<std::pair<std::string, BasicJson>> obj_;

    const BasicJson *find(const std::string &key) const {
        for (const auto &kv : obj_)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }

    template <class T, class Enable = void>
    struct Get;
    template <class T>
    struct Get<T, typename std::enable_if<std::is_arithmetic<T>::value && !std::is_same<T, bool>::value>::type> {
        static T from(const BasicJson &j) {
            if (j.kind_ != Kind::Number) throw std::runtime_error("json: not a number");
            return (T)j.num_;
        }
    };
    template <class T>
    struct Get<T, typename std::enable_if<std::is_same<T, bool>::value>::type> {
        static T from(const BasicJson &j) {
            if (j.kind_ != Kind::Bool) throw std::runtime_error("json: not a boolean");
            return j.b_;
        }
    };
    template <class T>
    struct Get<T, typename std::enable_if<std::is_same<T, std::string>::value>::type> {
        static T from(const BasicJson &j) {
            if (j.kind_ != Kind::String) throw std::runtime_error("json: not a string");
            return j.str_;
        }
    };

    static void quote(std::string &out, const std::string &s) {
        out += '"';
        for (unsigned char ch : s) {
            switch (ch) {
            case '"': out += "\\\""; break;
            case '\\': out += "\\\\"; break;
            case '\b': out += "\\b"; break;
            case '\f': out += "\\f"; break;
            case '\n': out += "\\n"; break;
            case '\r': out += "\\r"; break;
            case '\t': out += "\\t"; break;
            default:
                if (ch < 0x20) {
                    char buf[8];
                    std::snprintf(buf, sizeof(buf), "\\u%04x", ch);
                    out += buf;
                } else {
                    out += (char)ch;
                }
            }
        }
        out += '"';
    }
    // Numbers are written the way the reference's JSON library writes them:
    // integers as integers; a double as the digit string of the Grisu2
    // algorithm (F. Loitsch, "Printing floating-point numbers quickly and
    // accurately with integers", PLDI 2010: 64-bit scaled boundaries narrowed
    // by one unit each side, digits of the upper boundary until the
    // remainder fits the interval, last digit moved toward the value), laid
    // out as "ddd.0", "d.dd" or "0.000dd" when the decimal point falls within
    // [-4, 15] digits of the first digit, else "d.dde+XX" with a signed
    // exponent of at least two digits.  The text always reads back to the
    // same double; it is usually, not always, the shortest such text.
    static void number(std::string &out, double v, bool integral) {
        char buf[48];
        if (integral && std::fabs(v) < 9.0e15) {
            std::snprintf(buf, sizeof(buf), "%lld", (long long)v);
            out += buf;
            return;
        }
        if (!std::isfinite(v)) {
            out += "null";
            return;
        }
        if (std::signbit(v)) {
            out += '-';
            v = -v;
        }
        if (v == 0.0) {
            out += "0.0";
            return;
        }
        int e10 = 0;
        const std::string dig = Grisu2::digits(v, e10);  // v = dig * 10^e10
        const int k = (int)dig.size(), n = k + e10;     // v = 0.dig * 10^n
        if (k <= n && n <= 15) {
            out += dig;
            out.append((size_t)(n - k), '0');
            out += ".0";
        } else if (0 < n && n <= 15) {
            out += dig.substr(0, (size_t)n);
            out += '.';
            out += dig.substr((size_t)n);
        } else if (-4 < n && n <= 0) {
            out += "0.";
            out.append((size_t)(-n), '0');
            out += dig;
        } else {
            out += dig[0];
            if (k > 1) {
                out += '.';
                out += dig.substr(1);
            }
            std::snprintf(buf, sizeof(buf), "e%c%02d", n - 1 < 0 ? '-' : '+', std::abs(n - 1));
            out += buf;
        }
    }

    struct Grisu2 {
        struct Fp {
            uint64_t f;
            int e;
        };
        // (x * y + 2^63) >> 64: the product's upper half, rounded
        static Fp mul(Fp x, Fp y) {
            const unsigned __int128 p = (unsigned __int128)x.f * y.f + ((unsigned __int128)1 << 63);
            return Fp{(uint64_t)(p >> 64), x.e + y.e + 64};
        }
        static Fp normalize(Fp x) {
            while (!(x.f >> 63)) {
                x.f <<= 1;
                x.e--;
            }
            return x;
        }
        // 10^k for k = -300, -292, ..., 324 as f * 2^e with 2^63 <= f < 2^64,
        // f rounded to nearest (exact rational arithmetic)
        struct Pow10 {
            uint64_t f;
            int e, k;
        };
        static const Pow10 &cached(int binary_e) {
            static const Pow10 table[79] = {
            {0xAB70FE17C79AC6CAull, -1060, -300},
            {0xFF77B1FCBEBCDC4Full, -1034, -292},
            {0xBE5691EF416BD60Cull, -1007, -284},
            {0x8DD01FAD907FFC3Cull, -980, -276},
            {0xD3515C2831559A83ull, -954, -268},
            {0x9D71AC8FADA6C9B5ull, -927, -260},
            {0xEA9C227723EE8BCBull, -901, -252},
            {0xAECC49914078536Dull, -874, -244},
            {0x823C12795DB6CE57ull, -847, -236},
            {0xC21094364DFB5637ull, -821, -228},
            {0x9096EA6F3848984Full, -794, -220},
            {0xD77485CB25823AC7ull, -768, -212},
            {0xA086CFCD97BF97F4ull, -741, -204},
            {0xEF340A98172AACE5ull, -715, -196},
            {0xB23867FB2A35B28Eull, -688, -188},
            {0x84C8D4DFD2C63F3Bull, -661, -180},
            {0xC5DD44271AD3CDBAull, -635, -172},
            {0x936B9FCEBB25C996ull, -608, -164},
            {0xDBAC6C247D62A584ull, -582, -156},
            {0xA3AB66580D5FDAF6ull, -555, -148},
            {0xF3E2F893DEC3F126ull, -529, -140},
            {0xB5B5ADA8AAFF80B8ull, -502, -132},
            {0x87625F056C7C4A8Bull, -475, -124},
            {0xC9BCFF6034C13053ull, -449, -116},
            {0x964E858C91BA2655ull, -422, -108},
            {0xDFF9772470297EBDull, -396, -100},
            {0xA6DFBD9FB8E5B88Full, -369, -92},
            {0xF8A95FCF88747D94ull, -343, -84},
            {0xB94470938FA89BCFull, -316, -76},
            {0x8A08F0F8BF0F156Bull, -289, -68},
            {0xCDB02555653131B6ull, -263, -60},
            {0x993FE2C6D07B7FACull, -236, -52},
            {0xE45C10C42A2B3B06ull, -210, -44},
            {0xAA242499697392D3ull, -183, -36},
            {0xFD87B5F28300CA0Eull, -157, -28},
            {0xBCE5086492111AEBull, -130, -20},
            {0x8CBCCC096F5088CCull, -103, -12},
            {0xD1B71758E219652Cull, -77, -4},
            {0x9C40000000000000ull, -50, 4},
            {0xE8D4A51000000000ull, -24, 12},
            {0xAD78EBC5AC620000ull, 3, 20},
            {0x813F3978F8940984ull, 30, 28},
            {0xC097CE7BC90715B3ull, 56, 36},
            {0x8F7E32CE7BEA5C70ull, 83, 44},
            {0xD5D238A4ABE98068ull, 109, 52},
            {0x9F4F2726179A2245ull, 136, 60},
            {0xED63A231D4C4FB27ull, 162, 68},
            {0xB0DE65388CC8ADA8ull, 189, 76},
            {0x83C7088E1AAB65DBull, 216, 84},
            {0xC45D1DF942711D9Aull, 242, 92},
            {0x924D692CA61BE758ull, 269, 100},
            {0xDA01EE641A708DEAull, 295, 108},
            {0xA26DA3999AEF774Aull, 322, 116},
            {0xF209787BB47D6B85ull, 348, 124},
            {0xB454E4A179DD1877ull, 375, 132},
            {0x865B86925B9BC5C2ull, 402, 140},
            {0xC83553C5C8965D3Dull, 428, 148},
            {0x952AB45CFA97A0B3ull, 455, 156},
            {0xDE469FBD99A05FE3ull, 481, 164},
            {0xA59BC234DB398C25ull, 508, 172},
            {0xF6C69A72A3989F5Cull, 534, 180},
            {0xB7DCBF5354E9BECEull, 561, 188},
            {0x88FCF317F22241E2ull, 588, 196},
            {0xCC20CE9BD35C78A5ull, 614, 204},
            {0x98165AF37B2153DFull, 641, 212},
            {0xE2A0B5DC971F303Aull, 667, 220},
            {0xA8D9D1535CE3B396ull, 694, 228},
            {0xFB9B7CD9A4A7443Cull, 720, 236},
            {0xBB764C4CA7A44410ull, 747, 244},
            {0x8BAB8EEFB6409C1Aull, 774, 252},
            {0xD01FEF10A657842Cull, 800, 260},
            {0x9B10A4E5E9913129ull, 827, 268},
            {0xE7109BFBA19C0C9Dull, 853, 276},
            {0xAC2820D9623BF429ull, 880, 284},
            {0x80444B5E7AA7CF85ull, 907, 292},
            {0xBF21E44003ACDD2Dull, 933, 300},
            {0x8E679C2F5E44FF8Full, 960, 308},
            {0xD433179D9C8CB841ull, 986, 316},
            {0x9E19DB92B4E31BA9ull, 1013, 324},
            };
            // the smallest k with 10^k * 2^binary_e scaled into [2^-60, 2^-32)
            const int f = -60 - binary_e - 1;
            const int k = (f * 78913) / (1 << 18) + (f > 0 ? 1 : 0);
            return table[(300 + k + 7) / 8];
        }
        static std::string digits(double v, int &e10) {
            uint64_t bits;
            std::memcpy(&bits, &v, sizeof(bits));
            const uint64_t F = bits & ((uint64_t(1) << 52) - 1);
            const int E = (int)(bits >> 52);
            const Fp w0 = E == 0 ? Fp{F, -1074} : Fp{F + (uint64_t(1) << 52), E - 1075};
            // boundaries: midpoints to the neighbouring doubles
            const bool closer_below = F == 0 && E > 1;
            const Fp mp = normalize(Fp{2 * w0.f + 1, w0.e - 1});
            Fp mm = closer_below ? Fp{4 * w0.f - 1, w0.e - 2} : Fp{2 * w0.f - 1, w0.e - 1};
            mm.f <<= (mm.e - mp.e);
            mm.e = mp.e;
            const Fp w = normalize(w0);
            const Pow10 &c = cached(mp.e);
            const Fp cf{c.f, c.e};
            const Fp sw = mul(w, cf), sm = mul(mm, cf), sp = mul(mp, cf);
            const Fp lo{sm.f + 1, sm.e}, hi{sp.f - 1, sp.e};
            e10 = -c.k;
            // digits of hi while the remainder exceeds the interval width
            uint64_t delta = hi.f - lo.f, dist = hi.f - sw.f;
            const int sh = -hi.e;
            const uint64_t one = uint64_t(1) << sh;
            uint32_t p1 = (uint32_t)(hi.f >> sh);
            uint64_t p2 = hi.f & (one - 1);
            std::string d;
            uint32_t pow10 = 1;
            int n = 1;
            while (pow10 <= p1 / 10) {
                pow10 *= 10;
                n++;
            }
            while (n > 0) {
                d += (char)('0' + p1 / pow10);
                p1 %= pow10;
                n--;
                const uint64_t rest = ((uint64_t)p1 << sh) + p2;
                if (rest <= delta) {
                    e10 += n;
                    round_toward(d, dist, delta, rest, (uint64_t)pow10 << sh);
                    return d;
                }
                pow10 /= 10;
            }
            int m = 0;
            for (;;) {
                p2 *= 10;
                d += (char)('0' + (p2 >> sh));
                p2 &= one - 1;
                m++;
                delta *= 10;
                dist *= 10;
                if (p2 <= delta) break;
            }
            e10 -= m;
            round_toward(d, dist, delta, p2, one);
            return d;
        }
        // lower the last digit while that brings the text closer to the value
        static void round_toward(std::string &d, uint64_t dist, uint64_t delta, uint64_t rest, uint64_t ten_k) {
            while (rest < dist && delta - rest >= ten_k && (rest + ten_k < dist || dist - rest > rest + ten_k - dist)) {
                d.back()--;
                rest += ten_k;
            }
        }
    };
    void write(std::string &out, int indent, int level) const {
        auto nl = [&](int lv) {
            if (indent >= 0) {
                out += '\n';
                out.append((size_t)(indent * lv), ' ');
            }
        };
        switch (kind_) {
        case Kind::Null: out += "null"; break;
        case Kind::Bool: out += b_ ? "true" : "false"; break;
        case Kind::Number: number(out, num_, integral_); break;
        case Kind::String: quote(out, str_); break;
        case Kind::Array:
            if (arr_.empty()) {
                out += "[]";
                break;
            }
            out += '[';
            for (size_t q = 0; q < arr_.size(); q++) {
                if (q) out += ',';
                nl(level + 1);
                arr_[q].write(out, indent, level + 1);
            }
            nl(level);
            out += ']';
            break;
        case Kind::Object: {
            if (obj_.empty()) {
                out += "{}";
                break;
            }
            std::vector<const std::pair<std::string, BasicJson> *> order;
            for (const auto &kv : obj_) order.push_back(&kv);
            if (!ORDERED)
                std::sort(order.begin(), order.end(), [](const auto *a, const auto *b) { return a->first < b->first; });
            out += '{';
            for (size_t q = 0; q < order.size(); q++) {
                if (q) out += ',';
                nl(level + 1);
                quote(out, order[q]->first);
                out += indent >= 0 ? ": " : ":";
                order[q]->second.write(out, indent, level + 1);
            }
            nl(level);
            out += '}';
            break;
        }
        }
    }

    struct Parser {
        const char *s, *e;
        [[noreturn]] void fail(const char *what) const {
            throw std::runtime_error(std::string("json parse error: ") + what);
        }
        void ws() {
            while (s < e && (*s == ' ' || *s == '\t' || *s == '\n' || *s == '\r')) s++;
        }
        bool lit(const char *w) {
            const size_t n = std::strlen(w);
            if ((size_t)(e - s) >= n && std::strncmp(s, w, n) == 0) {
                s += n;
                return true;
            }
            return false;
        }
        BasicJson value() {
            if (s >= e) fail("unexpected end");
            switch (*s) {
            case '{': return object_();
            case '[': return array_();
            case '"': return BasicJson(string_());
            case 't':
                if (lit("true")) return BasicJson(true);
                fail("bad literal");
            case 'f':
                if (lit("false")) return BasicJson(false);
                fail("bad literal");
            case 'n':
                if (lit("null")) return BasicJson();
                fail("bad literal");
            default: return number_();
            }
        }
        BasicJson number_() {
            const char *b = s;
            bool integral = true;
            if (s < e && (*s == '-' || *s == '+')) s++;
            while (s < e && ((*s >= '0' && *s <= '9') || *s == '.' || *s == 'e' || *s == 'E' || *s == '-' ||
                             *s == '+')) {
                if (*s == '.' || *s == 'e' || *s == 'E') integral = false;
                s++;
            }
            if (s == b) fail("bad value");
            const std::string tok(b, s);
            char *end = nullptr;
            const double v = std::strtod(tok.c_str(), &end);
            if (!end || *end) fail("bad number");
            BasicJson j(v);
            j.integral_ = integral;
            return j;
        }
        std::string string_() {
            if (*s != '"') fail("expected a string");
            s++;
            std::string out;
            while (s < e && *s != '"') {
                if (*s == '\\') {
                    if (++s >= e) fail("bad escape");
                    switch (*s) {
                    case '"': out += '"'; break;
                    case '\\': out += '\\'; break;
                    case '/': out += '/'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'n': out += '\n'; break;
                    case 'r': out += '\r'; break;
                    case 't': out += '\t'; break;
                    case 'u': {
                        if (e - s < 5) fail("bad \\u escape");
                        const unsigned cp = (unsigned)std::strtoul(std::string(s + 1, s + 5).c_str(), nullptr, 16);
                        s += 4;
                        if (cp < 0x80) {
                            out += (char)cp;
                        } else if (cp < 0x800) {
                            out += (char)(0xc0 | (cp >> 6));
                            out += (char)(0x80 | (cp & 0x3f));
                        } else {
                            out += (char)(0xe0 | (cp >> 12));
                            out += (char)(0x80 | ((cp >> 6) & 0x3f));
                            out += (char)(0x80 | (cp & 0x3f));
                        }
                        break;
                    }
                    default: fail("bad escape");
                    }
                    s++;
                } else {
                    out += *s++;
                }
            }
            if (s >= e) fail("unterminated string");
            s++;
            return out;
        }
        BasicJson array_() {
            s++;
            BasicJson j = BasicJson::array();
            ws();
            if (s < e && *s == ']') {
                s++;
                return j;
            }
            for (;;) {
                ws();
                j.arr_.push_back(value());
                ws();
                if (s < e && *s == ',') {
                    s++;
                    continue;
                }
                if (s < e && *s == ']') {
                    s++;
                    return j;
                }
                fail("expected , or ]");
            }
        }
        BasicJson object_() {
            s++;
            BasicJson j = BasicJson::object();
            ws();
            if (s < e && *s == '}') {
                s++;
                return j;
            }
            for (;;) {
                ws();
                std::string key = string_();
                ws();
                if (s >= e || *s != ':') fail("expected :");
                s++;
                ws();
                BasicJson v = value();
                bool found = false;
                for (auto &kv : j.obj_)
                    if (kv.first == key) {  // a repeated key keeps the last value
                        kv.second = v;
                        found = true;
                    }
                if (!found) j.obj_.emplace_back(std::move(key), std::move(v));
                ws();
                if (s < e && *s == ',') {
                    s++;
                    continue;
                }
                if (s < e && *s == '}') {
                    s++;
                    return j;
                }
                fail("expected , or }");
            }
        }
    };
};

typedef BasicJson<false> Json;
typedef BasicJson<true> OrderedJson;

}  // namespace Falm

#endif
