#include "ParmParse.H"

#include <fstream>
#include <sstream>

namespace bamrex {

std::map<std::string, std::vector<std::string>> &ParmParse::table()
{
    static std::map<std::string, std::vector<std::string>> t;
    return t;
}

static std::vector<std::string> tokenize(const std::string &s)
{
    std::vector<std::string> out;
    std::string              cur;
    char                     quote = 0;
    for (char c : s)
    {
        if (quote)
        {
            if (c == quote)
                quote = 0;
            else
                cur += c;
        }
        else if (c == '"' || c == '\'')
            quote = c;
        else if (c == ' ' || c == '\t')
        {
            if (!cur.empty())
                out.push_back(cur);
            cur.clear();
        }
        else
            cur += c;
    }
    if (!cur.empty())
        out.push_back(cur);
    return out;
}

void ParmParse::add_line(const std::string &raw)
{
    std::string line = raw.substr(0, raw.find('#'));
    const auto  eq   = line.find('=');
    if (eq == std::string::npos)
        return;
    std::string key = line.substr(0, eq);
    key.erase(0, key.find_first_not_of(" \t"));
    key.erase(key.find_last_not_of(" \t") + 1);
    auto vals = tokenize(line.substr(eq + 1));
    if (!key.empty() && !vals.empty())
        table()[key] = vals;
}

void ParmParse::addfile(const std::string &path)
{
    std::ifstream f(path);
    if (!f)
        throw std::runtime_error("ParmParse: cannot open input file " + path);
    std::string line;
    while (std::getline(f, line))
        add_line(line);
}

void ParmParse::Initialize(int argc, char **argv)
{
    if (argc > 1)
        addfile(argv[1]);
    for (int i = 2; i < argc; ++i)
        add_line(argv[i]);
}

std::string ParmParse::full(const std::string &k) const
{
    return m_prefix.empty() ? k : m_prefix + "." + k;
}

bool ParmParse::contains(const std::string &key) const
{
    return table().count(full(key)) > 0;
}

template <class T> static T convert(const std::string &s)
{
    std::istringstream is(s);
    T                  v;
    is >> v;
    if (is.fail())
        throw std::runtime_error("ParmParse: cannot parse value '" + s + "'");
    return v;
}
template <> std::string convert<std::string>(const std::string &s) { return s; }
template <> bool convert<bool>(const std::string &s)
{
    if (s == "true" || s == "1")
        return true;
    if (s == "false" || s == "0")
        return false;
    throw std::runtime_error("ParmParse: not a bool '" + s + "'");
}

template <class T> bool ParmParse::query(const std::string &key, T &v) const
{
    auto it = table().find(full(key));
    if (it == table().end())
        return false;
    v = convert<T>(it->second.at(0));
    return true;
}

template <class T> void ParmParse::get(const std::string &key, T &v) const
{
    if (!query(key, v))
        throw std::runtime_error("ParmParse::get: required key " + full(key)
                                 + " not found");
}

template <class T>
bool ParmParse::queryarr(const std::string &key, std::vector<T> &v) const
{
    auto it = table().find(full(key));
    if (it == table().end())
        return false;
    v.clear();
    for (auto &s : it->second)
        v.push_back(convert<T>(s));
    return true;
}

template <class T>
void ParmParse::getarr(const std::string &key, std::vector<T> &v) const
{
    if (!queryarr(key, v))
        throw std::runtime_error("ParmParse::getarr: required key " + full(key)
                                 + " not found");
}

template bool ParmParse::query<int>(const std::string &, int &) const;
template bool ParmParse::query<double>(const std::string &, double &) const;
template bool ParmParse::query<bool>(const std::string &, bool &) const;
template bool ParmParse::query<std::string>(const std::string &, std::string &) const;
template void ParmParse::get<int>(const std::string &, int &) const;
template void ParmParse::get<double>(const std::string &, double &) const;
template void ParmParse::get<std::string>(const std::string &, std::string &) const;
template void ParmParse::get<bool>(const std::string &, bool &) const;
template bool ParmParse::queryarr<int>(const std::string &, std::vector<int> &) const;
template bool ParmParse::queryarr<double>(const std::string &, std::vector<double> &) const;
template bool ParmParse::queryarr<std::string>(const std::string &,
                                               std::vector<std::string> &) const;
template void ParmParse::getarr<int>(const std::string &, std::vector<int> &) const;
template void ParmParse::getarr<double>(const std::string &, std::vector<double> &) const;

} // namespace bamrex
