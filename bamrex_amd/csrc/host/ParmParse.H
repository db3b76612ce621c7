// Minimal amrex::ParmParse work-alike for the reference's input files:
// `prefix.key = v1 v2 ...  # comment` lines, later definitions win, and
// command-line `key=value` overrides after the input file name
// (the way the reference's run scripts pass amr.n_cell='1024 1024').
#pragma once
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace bamrex {

class ParmParse
{
  public:
    explicit ParmParse(std::string prefix = "") : m_prefix(std::move(prefix)) {}

    static void Initialize(int argc, char **argv); // argv[1] = input file
    static void addfile(const std::string &path);
    static void add_line(const std::string &line);

    bool contains(const std::string &key) const;

    template <class T> void get(const std::string &key, T &v) const;
    template <class T> bool query(const std::string &key, T &v) const;
    template <class T> void getarr(const std::string &key, std::vector<T> &v) const;
    template <class T> bool queryarr(const std::string &key, std::vector<T> &v) const;

  private:
    std::string                                      full(const std::string &k) const;
    static std::map<std::string, std::vector<std::string>> &table();
    std::string                                      m_prefix;
};

} // namespace bamrex
