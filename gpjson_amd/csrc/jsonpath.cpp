// jsonpath.cpp -- JSONPath subset -> IR bytes.
//
// Restates the reference compiler: scanner J/jsonpath/JSONPathScanner.java:5-96,
// recursive-descent parser J/jsonpath/JSONPathParser.java:20-261, IR builder
// J/jsonpath/IRBuilder.java:19-81 and byte encoding
// J/jsonpath/IRByteOutputBuffer.java:40-69 (opcode = enum ordinal, LEB128
// varints, strings as varint length + UTF-8).  (J/ = gpjson/src/main/java/
// com/koenv/gpjson/.)  Error classes match: UnsupportedJSONPathException ->
// kUnsupported, any other JSONPathException -> kParseError.
#include "jsonpath.h"

#include <stdexcept>

namespace gj {
namespace {

struct PathError {
  bool unsupported;
  std::string msg;
};

class Scanner {
 public:
  explicit Scanner(const std::string& s) : s_(s) {}
  bool has_next() const { return pos_ + 1 < (int)s_.size(); }
  char next() {
    if (!has_next()) throw PathError{false, "Expected character, got EOF at " + std::to_string(pos_)};
    return s_[++pos_];
  }
  char peek() const {
    if (!has_next()) throw PathError{false, "Expected character, got EOF at " + std::to_string(pos_)};
    return s_[pos_ + 1];
  }
  int position() const { return pos_; }
  // Java substring(start + 1, end + 1)
  std::string substring(int start, int end) const { return s_.substr(start + 1, end - start); }
  void expect(char c) {
    char n = next();
    if (n != c)
      throw PathError{false, std::string("Expected character '") + c + "', got character '" + n + "' at " +
                                 std::to_string(pos_)};
  }
  bool skip_if(char c) {
    if (!has_next()) return false;
    if (peek() == c) { pos_++; return true; }
    return false;
  }
  PathError error(const std::string& m) const { return PathError{false, m + " at " + std::to_string(pos_)}; }
  PathError error_next(const std::string& m) const { return PathError{false, m + " at " + std::to_string(pos_ + 1)}; }
  PathError unsupported_next(const std::string& m) const {
    return PathError{true, m + " at " + std::to_string(pos_ + 1)};
  }
  void mark() { marks_.push_back(pos_); }
  void reset() { pos_ = marks_.back(); marks_.pop_back(); }

 private:
  const std::string& s_;
  int pos_ = -1;
  std::vector<int> marks_;
};

class IR {
 public:
  void op(uint8_t o) { buf.push_back(o); }
  void varint(uint32_t i) {
    while (i & 0xFFFFFF80u) {
      buf.push_back((uint8_t)((i & 0x7F) | 0x80));
      i >>= 7;
    }
    buf.push_back((uint8_t)(i & 0x7F));
  }
  void str(const std::string& s) {
    varint((uint32_t)s.size());
    buf.insert(buf.end(), s.begin(), s.end());
  }
  void property(const std::string& name) { op(4); str(name); }
  void index(int i) { op(5); varint((uint32_t)i); }
  void down() { op(3); level++; }
  void up() { op(2); level--; }
  void store() { op(1); results++; }
  void end() { op(0); }
  void string_equals(const std::string& s) { op(6); str("\"" + s + "\""); }

  std::vector<uint8_t> buf;
  int level = 0;
  int results = 0;
};

class Parser {
 public:
  explicit Parser(const std::string& s) : sc_(s) {}

  void compile(CompiledPath* out) {
    sc_.expect('$');
    next_expression();
    ir_.store();
    ir_.end();
    out->ir = ir_.buf;
    out->max_depth = max_level_;
    out->num_results = ir_.results;
  }

 private:
  void next_expression() {
    char c = sc_.peek();
    if (c == '.') dot_expression();
    else if (c == '[') index_expression();
    else throw sc_.unsupported_next("Unsupported expression type");
  }

  void dot_expression() {
    sc_.expect('.');
    if (sc_.peek() == '.') throw sc_.unsupported_next("Unsupported recursive descent");
    std::string prop = read_property();
    if (prop.empty()) throw sc_.error("Unexpected empty property");
    property_ir(prop);
    if (sc_.has_next()) next_expression();
  }

  void index_expression() {
    sc_.expect('[');
    char c = sc_.peek();
    if (c == '\'' || c == '"') {
      std::string prop = read_quoted();
      if (prop.empty()) throw sc_.error("Unexpected empty property");
      property_ir(prop);
    } else if (c >= '0' && c <= '9') {
      int idx = read_integer([](char e) { return e == ']' || e == ':' || e == ','; });
      switch (sc_.peek()) {
        case ':':
          range_expression(idx);
          return;
        case ',':
          throw sc_.unsupported_next("Unsupported multiple index expression");
        case ']':
          index_ir(idx);
          break;
        default:
          break;
      }
    } else if (c == '*') {
      throw sc_.unsupported_next("Unsupported wildcard expression");
    } else if (c == '?') {
      filter_expression();
    } else {
      throw sc_.error_next("Unexpected character in index, expected ', \", or an integer");
    }
    sc_.expect(']');
    if (sc_.has_next()) next_expression();
  }

  void range_expression(int start) {
    sc_.expect(':');
    int end = read_integer([](char e) { return e == ']'; });
    sc_.expect(']');
    int max_max = max_level_;
    for (int i = start; i < end; i++) {
      int start_level = ir_.level;
      ir_.index(i);
      ir_.down();
      sc_.mark();
      int cur_max = max_level_;
      if (sc_.has_next()) next_expression();
      if (max_level_ > max_max) max_max = max_level_;
      max_level_ = cur_max;
      if (i == end - 1) break;
      ir_.store();
      sc_.reset();
      int end_level = ir_.level;
      for (int j = 0; j < end_level - start_level; j++) ir_.up();
    }
    max_level_ = max_max + 1;
  }

  void filter_expression() {
    sc_.expect('?');
    sc_.expect('(');
    sc_.expect('@');
    while (sc_.skip_if(' ')) {
    }
    if (sc_.peek() == '=') {
      sc_.expect('=');
      sc_.expect('=');
      while (sc_.skip_if(' ')) {
      }
      std::string eq = read_quoted();
      ir_.string_equals(eq);
    } else {
      throw sc_.unsupported_next("Unsupported character for expression");
    }
    sc_.expect(')');
  }

  void property_ir(const std::string& name) {
    ir_.property(name);
    ir_.down();
    max_level_++;
  }

  void index_ir(int i) {
    ir_.index(i);
    ir_.down();
    max_level_++;
  }

  std::string read_property() {
    int start = sc_.position();
    while (sc_.has_next()) {
      char c = sc_.peek();
      if (c == ' ') throw sc_.error_next("Unexpected space");
      if (c == '.' || c == '[') break;
      sc_.next();
    }
    return sc_.substring(start, sc_.position());
  }

  std::string read_quoted() {
    char q = sc_.next();
    if (q != '\'' && q != '"') throw sc_.error("Invalid quoted string");
    int start = sc_.position();
    bool escaped = false;
    while (sc_.has_next()) {
      char c = sc_.peek();
      if (escaped) escaped = false;
      else if (c == '\\') escaped = true;
      else if (c == q) break;
      sc_.next();
    }
    int end = sc_.position();
    sc_.expect(q);
    return sc_.substring(start, end);
  }

  template <class Pred>
  int read_integer(Pred is_end) {
    int start = sc_.position();
    while (sc_.has_next()) {
      char c = sc_.peek();
      if (c >= '0' && c <= '9') { sc_.next(); continue; }
      if (is_end(c)) break;
      throw sc_.error("Invalid integer");
    }
    std::string s = sc_.substring(start, sc_.position());
    // Integer.parseInt: empty or out-of-range strings are rejected
    if (s.empty() || s.size() > 10) throw sc_.error("Invalid integer");
    long long v = std::stoll(s);
    if (v > 2147483647LL) throw sc_.error("Invalid integer");
    return (int)v;
  }

  Scanner sc_;
  IR ir_;
  int max_level_ = 0;
};

}  // namespace

PathStatus compile_path(const std::string& path, CompiledPath* out, std::string* err) {
  try {
    Parser p(path);
    p.compile(out);
    return PathStatus::kOk;
  } catch (const PathError& e) {
    if (err) *err = e.msg;
    return e.unsupported ? PathStatus::kUnsupported : PathStatus::kParseError;
  }
}

}  // namespace gj
