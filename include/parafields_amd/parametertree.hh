// parametertree.hh -- string-keyed configuration with typed access and INI
// parsing: the subset of Dune::ParameterTree / ParameterTreeParser that
// parafields-core reads (randomfield.hh:78-84, fieldtraits.hh:169-209).
// Vector-valued keys are whitespace separated ("grid.cells = 256 256").
#pragma once

#include <array>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace parafields_amd {

class ParameterTree
{
  std::map<std::string, std::string> values_;

  template <typename T>
  static T parse_one(const std::string& s)
  {
    if constexpr (std::is_same_v<T, std::string>)
      return s;
    else if constexpr (std::is_same_v<T, bool>) {
      if (s == "true" || s == "1" || s == "yes")
        return true;
      if (s == "false" || s == "0" || s == "no")
        return false;
      throw std::runtime_error("cannot parse bool from '" + s + "'");
    } else {
      std::istringstream is(s);
      T v;
      is >> v;
      if (is.fail())
        throw std::runtime_error("cannot parse value from '" + s + "'");
      return v;
    }
  }

  template <typename T>
  struct is_array : std::false_type
  {};
  template <typename U, std::size_t N>
  struct is_array<std::array<U, N>> : std::true_type
  {};
  template <typename T>
  struct is_vector : std::false_type
  {};
  template <typename U>
  struct is_vector<std::vector<U>> : std::true_type
  {};

public:
  std::string& operator[](const std::string& key) { return values_[key]; }

  bool hasKey(const std::string& key) const { return values_.count(key) > 0; }

  template <typename T>
  T get(const std::string& key) const
  {
    auto it = values_.find(key);
    if (it == values_.end())
      throw std::runtime_error("missing configuration key " + key);
    if constexpr (is_array<T>::value) {
      T out{};
      std::istringstream is(it->second);
      std::string tok;
      std::size_t i = 0;
      while (is >> tok) {
        if (i >= out.size())
          throw std::runtime_error("too many values for key " + key);
        out[i++] = parse_one<typename T::value_type>(tok);
      }
      if (i != out.size())
        throw std::runtime_error("too few values for key " + key);
      return out;
    } else if constexpr (is_vector<T>::value) {
      T out;
      std::istringstream is(it->second);
      std::string tok;
      while (is >> tok)
        out.push_back(parse_one<typename T::value_type>(tok));
      return out;
    } else {
      std::string v = it->second;
      const auto b = v.find_first_not_of(" \t"), e = v.find_last_not_of(" \t");
      v = (b == std::string::npos) ? "" : v.substr(b, e - b + 1);
      return parse_one<T>(v);
    }
  }

  template <typename T>
  T get(const std::string& key, const T& def) const
  {
    return hasKey(key) ? get<T>(key) : def;
  }

  // INI: "[section]" prefixes keys with "section."; '#' starts a comment.
  void readINI(std::istream& in)
  {
    std::string line, section;
    while (std::getline(in, line)) {
      const auto hash = line.find('#');
      if (hash != std::string::npos)
        line = line.substr(0, hash);
      const auto b = line.find_first_not_of(" \t\r");
      if (b == std::string::npos)
        continue;
      line = line.substr(b, line.find_last_not_of(" \t\r") - b + 1);
      if (line.front() == '[' && line.back() == ']') {
        section = line.substr(1, line.size() - 2);
        const auto sb = section.find_first_not_of(" \t"), se = section.find_last_not_of(" \t");
        section = sb == std::string::npos ? std::string() : section.substr(sb, se - sb + 1);
        continue;
      }
      const auto eq = line.find('=');
      if (eq == std::string::npos)
        throw std::runtime_error("malformed configuration line: " + line);
      auto trim = [](std::string s) {
        const auto bb = s.find_first_not_of(" \t"), ee = s.find_last_not_of(" \t");
        return bb == std::string::npos ? std::string() : s.substr(bb, ee - bb + 1);
      };
      std::string key = trim(line.substr(0, eq)), val = trim(line.substr(eq + 1));
      if (val.size() >= 2 && val.front() == '"' && val.back() == '"')
        val = val.substr(1, val.size() - 2); // quoted, as report() writes
      if (!section.empty())
        key = section + "." + key;
      values_[key] = val;
    }
  }

  // Dune::ParameterTree::report layout: the keys without a section as
  // key = "value", then each section as [ name ] followed by its keys.
  // readINI (and Dune's INI parser) read it back.
  void report(std::ostream& out) const
  {
    std::map<std::string, std::map<std::string, std::string>> sections;
    for (const auto& kv : values_) {
      const auto dot = kv.first.rfind('.');
      if (dot == std::string::npos)
        out << kv.first << " = \"" << kv.second << "\"" << std::endl;
      else
        sections[kv.first.substr(0, dot)][kv.first.substr(dot + 1)] = kv.second;
    }
    for (const auto& sec : sections) {
      out << "[ " << sec.first << " ]" << std::endl;
      for (const auto& kv : sec.second)
        out << kv.first << " = \"" << kv.second << "\"" << std::endl;
    }
  }

  void readINIFile(const std::string& file)
  {
    std::ifstream in(file);
    if (!in)
      throw std::runtime_error("cannot open " + file);
    readINI(in);
  }
};

} // namespace parafields_amd
