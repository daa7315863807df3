// YAML-subset reader (see yaml.hpp for the supported subset).
#include "yaml.hpp"

#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <limits>
#include <map>
#include <sstream>

namespace frehg {
namespace yaml {

const Node* Node::get(const std::string& key) const {
  if (kind != Map) return nullptr;
  for (auto it = map.rbegin(); it != map.rend(); ++it)
    if (it->first == key) return it->second.get();
  return nullptr;
}

void Node::set(const std::string& key, std::shared_ptr<Node> value) {
  if (kind == Null) kind = Map;
  if (kind != Map) throw Error("yaml: set('" + key + "') on a node that is not a mapping");
  for (auto& kv : map)
    if (kv.first == key) {
      kv.second = std::move(value);
      return;
    }
  map.emplace_back(key, std::move(value));
}

namespace {

std::string where(const Node& n) { return n.line > 0 ? " (line " + std::to_string(n.line) + ")" : ""; }

const std::string& scalar_text(const Node& n, const char* what) {
  if (n.kind != Node::Scalar) throw Error(std::string("yaml: expected a scalar ") + what + where(n));
  return n.text;
}

}  // namespace

double Node::as_double() const {
  const std::string& t = scalar_text(*this, "number");
  if (t == ".inf" || t == ".Inf" || t == ".INF" || t == "+.inf" || t == "+.Inf" || t == "+.INF")
    return std::numeric_limits<double>::infinity();
  if (t == "-.inf" || t == "-.Inf" || t == "-.INF") return -std::numeric_limits<double>::infinity();
  if (t == ".nan" || t == ".NaN" || t == ".NAN") return std::numeric_limits<double>::quiet_NaN();
  if (t.empty()) throw Error("yaml: empty value where a number is expected" + where(*this));
  char* end = nullptr;
  errno = 0;
  const double v = std::strtod(t.c_str(), &end);
  if (end == t.c_str() || *end != '\0' || std::isspace((unsigned char)t[0]))
    throw Error("yaml: '" + t + "' is not a number" + where(*this));
  return v;
}

long long Node::as_long() const {
  const std::string& t = scalar_text(*this, "integer");
  char* end = nullptr;
  errno = 0;
  const long long v = std::strtoll(t.c_str(), &end, 10);
  if (t.empty() || end == t.c_str() || *end != '\0' || errno == ERANGE || std::isspace((unsigned char)t[0]))
    throw Error("yaml: '" + t + "' is not an integer" + where(*this));
  return v;
}

int Node::as_int() const {
  const long long v = as_long();
  if (v < std::numeric_limits<int>::min() || v > std::numeric_limits<int>::max())
    throw Error("yaml: '" + text + "' is out of the int range" + where(*this));
  return (int)v;
}

bool Node::as_bool() const {
  const std::string& t = scalar_text(*this, "boolean");
  static const char* yes[] = {"true", "True", "TRUE", "yes", "Yes", "YES", "y", "Y", "on", "On", "ON"};
  static const char* no[] = {"false", "False", "FALSE", "no", "No", "NO", "n", "N", "off", "Off", "OFF"};
  for (const char* w : yes)
    if (t == w) return true;
  for (const char* w : no)
    if (t == w) return false;
  throw Error("yaml: '" + t + "' is not a boolean" + where(*this));
}

const std::string& Node::as_string() const { return scalar_text(*this, "string"); }

std::shared_ptr<Node> scalar(const std::string& text) {
  auto n = std::make_shared<Node>();
  n->kind = Node::Scalar;
  n->text = text;
  return n;
}

namespace {

struct Line {
  int no;
  int indent;
  std::string s;  // content: indentation, comment and trailing blanks removed
};

bool is_blank(char c) { return c == ' ' || c == '\t'; }

std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && is_blank(s[a])) ++a;
  while (b > a && (is_blank(s[b - 1]) || s[b - 1] == '\r')) --b;
  return s.substr(a, b - a);
}

bool plain_null(const std::string& t) { return t.empty() || t == "~" || t == "null" || t == "Null" || t == "NULL"; }

// A quote opens a quoted scalar only at the start of a value (after "key: ", "- ", "[", "{" or
// ","); an apostrophe inside a plain scalar does not.
bool quote_opens(const std::string& s, size_t i) {
  size_t j = i;
  while (j > 0 && is_blank(s[j - 1])) --j;
  if (j == 0) return true;
  const char p = s[j - 1];
  return p == ':' || p == '-' || p == '[' || p == '{' || p == ',';
}

// Removes a comment: '#' at the start or after a blank, outside quotes.
std::string strip_comment(const std::string& s) {
  char q = 0;
  for (size_t i = 0; i < s.size(); ++i) {
    const char c = s[i];
    if (q) {
      if (q == '\'' && c == '\'') {
        if (i + 1 < s.size() && s[i + 1] == '\'') {
          ++i;
          continue;
        }
        q = 0;
      } else if (q == '"' && c == '\\') {
        ++i;
      } else if (q == '"' && c == '"') {
        q = 0;
      }
      continue;
    }
    if ((c == '\'' || c == '"') && quote_opens(s, i)) {
      q = c;
      continue;
    }
    if (c == '#' && (i == 0 || is_blank(s[i - 1]))) return s.substr(0, i);
  }
  return s;
}

class Parser {
 public:
  Parser(const std::string& text, std::string source) : src_(std::move(source)) {
    std::istringstream in(text);
    std::string raw;
    int no = 0;
    bool started = false;
    while (std::getline(in, raw)) {
      ++no;
      if (!raw.empty() && raw.back() == '\r') raw.pop_back();
      size_t ind = 0;
      while (ind < raw.size() && raw[ind] == ' ') ++ind;
      if (ind < raw.size() && raw[ind] == '\t') {
        const std::string rest = trim(strip_comment(raw.substr(ind)));
        if (!rest.empty()) fail(no, "tab in indentation");
      }
      const std::string s = trim(strip_comment(raw.substr(ind)));
      if (s.empty()) continue;
      if (ind == 0 && (s == "---" || s.rfind("--- ", 0) == 0)) {
        if (started) fail(no, "several documents in one file are not supported");
        if (s != "---") fail(no, "content after '---' is not supported");
        continue;
      }
      if (ind == 0 && s == "...") break;
      if (s[0] == '%') fail(no, "directives are not supported");
      started = true;
      lines_.push_back({no, (int)ind, s});
    }
  }

  std::shared_ptr<Node> document() {
    if (lines_.empty()) return std::make_shared<Node>();
    auto n = block(-1);
    if (pos_ < lines_.size()) fail(lines_[pos_].no, "unexpected content (check the indentation)");
    return n;
  }

 private:
  [[noreturn]] void fail(int line, const std::string& msg) const {
    throw Error("yaml: " + src_ + ":" + std::to_string(line) + ": " + msg);
  }

  static bool is_dash(const std::string& s) { return s == "-" || s.rfind("- ", 0) == 0; }

  // Position of the ':' that ends a mapping key ("key: value" or "key:"), or npos.
  static size_t key_colon(const std::string& s) {
    if (s.empty() || s[0] == '[' || s[0] == '{' || s[0] == '*' || s[0] == '&' || s[0] == '!' || s[0] == '|' ||
        s[0] == '>')
      return std::string::npos;
    size_t i = 0;
    if (s[0] == '\'' || s[0] == '"') {
      const char q = s[0];
      for (i = 1; i < s.size(); ++i) {
        if (q == '\'' && s[i] == '\'') {
          if (i + 1 < s.size() && s[i + 1] == '\'') {
            ++i;
            continue;
          }
          break;
        }
        if (q == '"' && s[i] == '\\') {
          ++i;
          continue;
        }
        if (q == '"' && s[i] == '"') break;
      }
      if (i >= s.size()) return std::string::npos;
      ++i;
      while (i < s.size() && is_blank(s[i])) ++i;
      return (i < s.size() && s[i] == ':' && (i + 1 == s.size() || is_blank(s[i + 1]))) ? i : std::string::npos;
    }
    for (i = 0; i < s.size(); ++i)
      if (s[i] == ':' && (i + 1 == s.size() || is_blank(s[i + 1]))) return i;
    return std::string::npos;
  }

  std::string key_text(const std::string& k, int line) const {
    const std::string t = trim(k);
    if (!t.empty() && (t[0] == '\'' || t[0] == '"')) {
      size_t i = 0;
      std::string out = quoted(t, i, line);
      if (i != t.size()) fail(line, "unexpected text after a quoted key");
      return out;
    }
    if (t.rfind("? ", 0) == 0 || t == "?") fail(line, "complex keys are not supported");
    return t;
  }

  // Parses a quoted scalar starting at s[i] (a quote); leaves i after the closing quote.
  std::string quoted(const std::string& s, size_t& i, int line) const {
    const char q = s[i++];
    std::string out;
    for (; i < s.size(); ++i) {
      const char c = s[i];
      if (q == '\'') {
        if (c == '\'') {
          if (i + 1 < s.size() && s[i + 1] == '\'') {
            out.push_back('\'');
            ++i;
            continue;
          }
          ++i;
          return out;
        }
        out.push_back(c);
      } else {
        if (c == '"') {
          ++i;
          return out;
        }
        if (c == '\\') {
          if (++i >= s.size()) break;
          switch (s[i]) {
            case 'n': out.push_back('\n'); break;
            case 't': out.push_back('\t'); break;
            case 'r': out.push_back('\r'); break;
            case '0': out.push_back('\0'); break;
            case '\\': out.push_back('\\'); break;
            case '"': out.push_back('"'); break;
            case '/': out.push_back('/'); break;
            case ' ': out.push_back(' '); break;
            default: fail(line, std::string("unsupported escape \\") + s[i]);
          }
          continue;
        }
        out.push_back(c);
      }
    }
    fail(line, "unterminated quoted scalar (multi-line quoted scalars are not supported)");
  }

  std::shared_ptr<Node> make_scalar(const std::string& t, bool q, int line) const {
    auto n = std::make_shared<Node>();
    n->line = line;
    if (!q && plain_null(t)) return n;
    n->kind = Node::Scalar;
    n->text = t;
    n->quoted = q;
    return n;
  }

  // A node made of the lines indented more than `parent` (or, for the root, from column 0).
  std::shared_ptr<Node> block(int parent) {
    const Line& L = lines_[pos_];
    if (L.indent <= parent) fail(L.no, "expected an indented block");
    if (is_dash(L.s)) return seq(L.indent);
    if (key_colon(L.s) != std::string::npos) return mapping(L.indent);
    // a lone value; a continuation line is left to the caller, which rejects it
    const int no = L.no, ind = L.indent;
    const std::string s = L.s;
    ++pos_;
    return inline_value(s, no, ind);
  }

  // Anchor prefix "&name" of a value: returns the name and strips it from `rest`.
  std::string take_anchor(std::string& rest, int line) const {
    if (rest.empty() || rest[0] != '&') return "";
    size_t e = 1;
    while (e < rest.size() && !is_blank(rest[e])) ++e;
    const std::string name = rest.substr(1, e - 1);
    if (name.empty()) fail(line, "empty anchor name");
    rest = trim(rest.substr(e));
    return name;
  }

  std::shared_ptr<Node> mapping(int ind) {
    auto node = std::make_shared<Node>();
    node->kind = Node::Map;
    node->line = lines_[pos_].no;
    while (pos_ < lines_.size() && lines_[pos_].indent == ind) {
      const Line L = lines_[pos_];
      if (is_dash(L.s)) fail(L.no, "a sequence item where a mapping key is expected");
      const size_t c = key_colon(L.s);
      if (c == std::string::npos) fail(L.no, "expected 'key: value'");
      const std::string key = key_text(L.s.substr(0, c), L.no);
      std::string rest = trim(L.s.substr(c + 1));
      ++pos_;
      const std::string anchor = take_anchor(rest, L.no);
      std::shared_ptr<Node> child;
      if (rest.empty()) {
        if (pos_ < lines_.size() && lines_[pos_].indent > ind)
          child = block(ind);
        else if (pos_ < lines_.size() && lines_[pos_].indent == ind && is_dash(lines_[pos_].s))
          child = seq(ind);  // "key:" followed by "- item" at the key's indentation
        else
          child = make_scalar("", false, L.no);
      } else {
        child = inline_value(rest, L.no, ind);
      }
      if (!anchor.empty()) anchors_[anchor] = child;
      node->set(key, child);
    }
    if (pos_ < lines_.size() && lines_[pos_].indent > ind)
      fail(lines_[pos_].no, "unexpected indentation");
    return node;
  }

  std::shared_ptr<Node> seq(int ind) {
    auto node = std::make_shared<Node>();
    node->kind = Node::Seq;
    node->line = lines_[pos_].no;
    while (pos_ < lines_.size() && lines_[pos_].indent == ind && is_dash(lines_[pos_].s)) {
      Line& L = lines_[pos_];
      const int no = L.no;
      size_t off = 1;
      while (off < L.s.size() && is_blank(L.s[off])) ++off;
      std::string rest = L.s.substr(off);
      const std::string anchor = take_anchor(rest, no);
      std::shared_ptr<Node> child;
      if (rest.empty()) {
        ++pos_;
        if (pos_ < lines_.size() && lines_[pos_].indent > ind)
          child = block(ind);
        else
          child = make_scalar("", false, no);
      } else if (anchor.empty() && (is_dash(rest) || key_colon(rest) != std::string::npos)) {
        // The item opens a block collection on the dash line: continue as if that text started
        // its own line at its column.
        L.indent = ind + (int)off;
        L.s = rest;
        child = is_dash(rest) ? seq(L.indent) : mapping(L.indent);
      } else {
        ++pos_;
        child = inline_value(rest, no, ind);
      }
      if (!anchor.empty()) anchors_[anchor] = child;
      node->items.push_back(child);
    }
    if (pos_ < lines_.size() && lines_[pos_].indent > ind)
      fail(lines_[pos_].no, "unexpected indentation");
    return node;
  }

  // The value after "key:" or "- " on one line; a flow collection may continue on the lines
  // that follow (they are consumed).
  std::shared_ptr<Node> inline_value(std::string rest, int line, int ind) {
    if (rest[0] == '*') {
      const std::string name = trim(rest.substr(1));
      auto it = anchors_.find(name);
      if (it == anchors_.end()) fail(line, "unknown alias *" + name);
      return it->second;
    }
    if (rest[0] == '!') fail(line, "tags are not supported");
    if (rest[0] == '|' || rest[0] == '>') fail(line, "block scalars (| and >) are not supported");
    if (rest[0] == '[' || rest[0] == '{') {
      while (!balanced(rest)) {
        if (pos_ >= lines_.size() || lines_[pos_].indent <= ind) fail(line, "unterminated flow collection");
        rest += " " + lines_[pos_].s;
        ++pos_;
      }
      size_t i = 0;
      auto n = flow(rest, i, line);
      skip_ws(rest, i);
      if (i != rest.size()) fail(line, "unexpected text after a flow collection");
      return n;
    }
    if (rest[0] == '\'' || rest[0] == '"') {
      size_t i = 0;
      const std::string t = quoted(rest, i, line);
      if (i != rest.size()) fail(line, "unexpected text after a quoted scalar");
      return make_scalar(t, true, line);
    }
    return make_scalar(rest, false, line);
  }

  static bool balanced(const std::string& s) {
    int depth = 0;
    char q = 0;
    for (size_t i = 0; i < s.size(); ++i) {
      const char c = s[i];
      if (q) {
        if (q == '\'' && c == '\'') {
          if (i + 1 < s.size() && s[i + 1] == '\'')
            ++i;
          else
            q = 0;
        } else if (q == '"' && c == '\\') {
          ++i;
        } else if (q == '"' && c == '"') {
          q = 0;
        }
        continue;
      }
      if ((c == '\'' || c == '"') && quote_opens(s, i)) q = c;
      else if (c == '[' || c == '{') ++depth;
      else if (c == ']' || c == '}') --depth;
    }
    return depth <= 0;
  }

  static void skip_ws(const std::string& s, size_t& i) {
    while (i < s.size() && is_blank(s[i])) ++i;
  }

  std::shared_ptr<Node> flow(const std::string& s, size_t& i, int line) {
    skip_ws(s, i);
    if (i >= s.size()) fail(line, "missing value in a flow collection");
    const char c = s[i];
    std::string anchor;
    if (c == '&') {
      size_t e = i + 1;
      while (e < s.size() && !is_blank(s[e]) && s[e] != ',' && s[e] != ']' && s[e] != '}') ++e;
      anchor = s.substr(i + 1, e - i - 1);
      i = e;
      auto n = flow(s, i, line);
      anchors_[anchor] = n;
      return n;
    }
    if (c == '*') {
      size_t e = i + 1;
      while (e < s.size() && !is_blank(s[e]) && s[e] != ',' && s[e] != ']' && s[e] != '}') ++e;
      const std::string name = s.substr(i + 1, e - i - 1);
      i = e;
      auto it = anchors_.find(name);
      if (it == anchors_.end()) fail(line, "unknown alias *" + name);
      return it->second;
    }
    if (c == '[') {
      auto n = std::make_shared<Node>();
      n->kind = Node::Seq;
      n->line = line;
      ++i;
      for (;;) {
        skip_ws(s, i);
        if (i >= s.size()) fail(line, "unterminated flow sequence");
        if (s[i] == ']') {
          ++i;
          return n;
        }
        n->items.push_back(flow(s, i, line));
        skip_ws(s, i);
        if (i < s.size() && s[i] == ',') {
          ++i;
          continue;
        }
        if (i < s.size() && s[i] == ']') {
          ++i;
          return n;
        }
        fail(line, "expected ',' or ']' in a flow sequence");
      }
    }
    if (c == '{') {
      auto n = std::make_shared<Node>();
      n->kind = Node::Map;
      n->line = line;
      ++i;
      for (;;) {
        skip_ws(s, i);
        if (i >= s.size()) fail(line, "unterminated flow mapping");
        if (s[i] == '}') {
          ++i;
          return n;
        }
        std::string key;
        if (s[i] == '\'' || s[i] == '"') {
          key = quoted(s, i, line);
        } else {
          size_t e = i;
          while (e < s.size() && s[e] != ':' && s[e] != ',' && s[e] != '}') ++e;
          key = trim(s.substr(i, e - i));
          i = e;
        }
        skip_ws(s, i);
        std::shared_ptr<Node> v;
        if (i < s.size() && s[i] == ':') {
          ++i;
          skip_ws(s, i);
          if (i < s.size() && (s[i] == ',' || s[i] == '}'))
            v = make_scalar("", false, line);
          else
            v = flow(s, i, line);
        } else {
          v = make_scalar("", false, line);
        }
        n->set(key, v);
        skip_ws(s, i);
        if (i < s.size() && s[i] == ',') {
          ++i;
          continue;
        }
        if (i < s.size() && s[i] == '}') {
          ++i;
          return n;
        }
        fail(line, "expected ',' or '}' in a flow mapping");
      }
    }
    if (c == '\'' || c == '"') return make_scalar(quoted(s, i, line), true, line);
    size_t e = i;
    while (e < s.size() && s[e] != ',' && s[e] != ']' && s[e] != '}') {
      if (s[e] == ':' && (e + 1 == s.size() || is_blank(s[e + 1])))
        fail(line, "a mapping inside a flow sequence needs braces");
      ++e;
    }
    const std::string t = trim(s.substr(i, e - i));
    i = e;
    return make_scalar(t, false, line);
  }

  std::string src_;
  std::vector<Line> lines_;
  size_t pos_ = 0;
  std::map<std::string, std::shared_ptr<Node>> anchors_;
};

void dump_into(const Node& n, int ind, std::ostringstream& out, bool inline_first) {
  const std::string pad(ind, ' ');
  auto scalar_out = [](const Node& s) -> std::string {
    if (s.kind == Node::Null) return "null";
    bool plain = !s.quoted && !s.text.empty();
    for (char c : s.text)
      if (c == ':' || c == '#' || c == '\'' || c == '"' || c == '[' || c == ']' || c == '{' || c == '}' ||
          c == ',' || c == '\n')
        plain = false;
    if (plain) return s.text;
    std::string q = "'";
    for (char c : s.text) {
      if (c == '\'') q += "''";
      else q.push_back(c);
    }
    return q + "'";
  };
  if (n.kind == Node::Map) {
    bool first = true;
    for (const auto& kv : n.map) {
      if (!(first && inline_first)) out << pad;
      first = false;
      out << kv.first << ":";
      const Node& v = *kv.second;
      if (v.kind == Node::Map && !v.map.empty()) {
        out << "\n";
        dump_into(v, ind + 2, out, false);
      } else if (v.kind == Node::Seq && !v.items.empty()) {
        out << "\n";
        dump_into(v, ind + 2, out, false);
      } else if (v.kind == Node::Map) {
        out << " {}\n";
      } else if (v.kind == Node::Seq) {
        out << " []\n";
      } else {
        out << " " << scalar_out(v) << "\n";
      }
    }
  } else if (n.kind == Node::Seq) {
    for (const auto& it : n.items) {
      out << pad << "- ";
      if (it->kind == Node::Map && !it->map.empty())
        dump_into(*it, ind + 2, out, true);
      else if (it->kind == Node::Seq && !it->items.empty()) {
        out << "\n";
        dump_into(*it, ind + 2, out, false);
      } else if (it->kind == Node::Map)
        out << "{}\n";
      else if (it->kind == Node::Seq)
        out << "[]\n";
      else
        out << scalar_out(*it) << "\n";
    }
  } else {
    out << pad << scalar_out(n) << "\n";
  }
}

}  // namespace

std::shared_ptr<Node> parse(const std::string& text, const std::string& source) {
  Parser p(text, source);
  return p.document();
}

std::shared_ptr<Node> load_file(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  if (!in) throw Error("yaml: cannot open '" + path + "'");
  std::ostringstream ss;
  ss << in.rdbuf();
  return parse(ss.str(), path);
}

std::string dump(const Node& n) {
  std::ostringstream out;
  dump_into(n, 0, out, false);
  return out.str();
}

}  // namespace yaml
}  // namespace frehg
