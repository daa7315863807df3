// A reader for the YAML subset the Frehg2 v2 configuration files use, standing in for yaml-cpp
// (which the reference links: src/io/Config.cpp:51-79, YAML::LoadFile / YAML::Load).
//
// Supported: block mappings and sequences by indentation (a sequence may sit at its key's own
// indentation, as PyYAML writes it), sequence items that open a mapping on the dash line,
// anchors (&name) and aliases (*name), flow sequences and mappings ([a, [b, c]], {k: v}, which
// may span lines), single- and double-quoted scalars, plain scalars, comments and document
// markers. Not supported (rejected with the line number): tags, block scalars (| and >),
// multi-line plain scalars, complex keys, several documents in one file.
//
// Conversions follow yaml-cpp's: as_double accepts decimal and exponent forms and .inf/.nan;
// as_int/as_long accept decimal integers only; as_bool accepts true/false, yes/no, y/n and
// on/off in lower, Title or UPPER case. A plain null, ~ or empty value is a null node.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace frehg {
namespace yaml {

class Error : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

struct Node {
  enum Kind { Null, Scalar, Seq, Map };
  Kind kind = Null;
  std::string text;     // scalar text (quotes removed, escapes resolved)
  bool quoted = false;  // a quoted scalar is never null
  int line = 0;         // 1-based source line (0 for nodes made by the program)
  std::vector<std::shared_ptr<Node>> items;                        // Seq
  std::vector<std::pair<std::string, std::shared_ptr<Node>>> map;  // Map, in file order

  bool is_null() const { return kind == Null; }
  bool is_scalar() const { return kind == Scalar; }
  bool is_seq() const { return kind == Seq; }
  bool is_map() const { return kind == Map; }
  size_t size() const { return kind == Seq ? items.size() : kind == Map ? map.size() : 0; }
  // Child of a mapping (the last one when a key repeats); null when absent or not a mapping.
  const Node* get(const std::string& key) const;
  const Node* at(size_t i) const { return (kind == Seq && i < items.size()) ? items[i].get() : nullptr; }
  // Set (or add) a mapping entry; the node becomes a mapping if it was null.
  void set(const std::string& key, std::shared_ptr<Node> value);

  double as_double() const;
  int as_int() const;
  long long as_long() const;
  bool as_bool() const;
  const std::string& as_string() const;
};

std::shared_ptr<Node> scalar(const std::string& text);

// Parses one document; `source` names it in error messages.
std::shared_ptr<Node> parse(const std::string& text, const std::string& source = "<string>");
std::shared_ptr<Node> load_file(const std::string& path);

// Block-style dump (for the resolved configuration in summaries and error reports).
std::string dump(const Node& n);

}  // namespace yaml
}  // namespace frehg
