// mpm_xml.cpp -- reader for the NairnMPM XML input (NairnMPM/input/NairnMPM.dtd), limited to the
// elements of the explicit-step path. It replaces the Xerces SAX2 pipeline of the reference
// (Common/System/CommonAnalysis.cpp:184-273 -> CommonReadHandler.cpp:50-340 ->
// Read_MPM/MPMReadHandler.cpp:83-1330, Read_MPM/Generators.cpp:89-1545) with a small in-memory
// parser: elements, attributes, text, comments, processing instructions and a DOCTYPE internal
// subset with <!ENTITY name "value"> declarations (as used by NairnMPM/input/XML_Input).
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "mpm_problem.h"

namespace {

struct XNode {
    std::string name;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::string text;
    std::vector<std::unique_ptr<XNode>> kids;
    const char *attr(const char *k) const {
        for (auto &a : attrs) if (a.first == k) return a.second.c_str();
        return nullptr;
    }
};

struct Parser {
    const char *s;
    size_t i = 0, n;
    std::map<std::string, std::string> entities;
    std::string err;
    explicit Parser(const char *t) : s(t), n(strlen(t)) {}

    bool starts(const char *p) const { return strncmp(s + i, p, strlen(p)) == 0; }
    void ws() { while (i < n && isspace((unsigned char)s[i])) i++; }
    bool fail(const std::string &m) {
        if (err.empty()) {
            int line = 1;
            for (size_t k = 0; k < i && k < n; k++) if (s[k] == '\n') line++;
            err = "XML input line " + std::to_string(line) + ": " + m;
        }
        return false;
    }
    std::string expand(const std::string &raw) {
        std::string out;
        for (size_t k = 0; k < raw.size(); k++) {
            if (raw[k] != '&') { out += raw[k]; continue; }
            size_t e = raw.find(';', k);
            if (e == std::string::npos) { out += raw[k]; continue; }
            std::string name = raw.substr(k + 1, e - k - 1);
            if (name == "lt") out += '<';
            else if (name == "gt") out += '>';
            else if (name == "amp") out += '&';
            else if (name == "quot") out += '"';
            else if (name == "apos") out += '\'';
            else if (!name.empty() && name[0] == '#') out += (char)strtol(name.c_str() + (name[1] == 'x' ? 2 : 1), nullptr, name[1] == 'x' ? 16 : 10);
            else if (entities.count(name)) out += expand(entities[name]);
            else { out += raw.substr(k, e - k + 1); }
            k = e;
        }
        return out;
    }
    bool skip_misc() {
        for (;;) {
            ws();
            if (starts("<?")) {
                size_t e = std::string(s + i).find("?>");
                if (e == std::string::npos) return fail("unterminated processing instruction");
                i += e + 2;
            } else if (starts("<!--")) {
                size_t e = std::string(s + i).find("-->");
                if (e == std::string::npos) return fail("unterminated comment");
                i += e + 3;
            } else if (starts("<!DOCTYPE")) {
                if (!doctype()) return false;
            } else return true;
        }
    }
    bool doctype() {
        i += 9;
        while (i < n && s[i] != '[' && s[i] != '>') {
            if (s[i] == '"' || s[i] == '\'') { char q = s[i++]; while (i < n && s[i] != q) i++; }
            i++;
        }
        if (i < n && s[i] == '[') {
            i++;
            for (;;) {
                ws();
                if (i >= n) return fail("unterminated DOCTYPE");
                if (s[i] == ']') { i++; break; }
                if (starts("<!--")) {
                    size_t e = std::string(s + i).find("-->");
                    if (e == std::string::npos) return fail("unterminated comment");
                    i += e + 3;
                } else if (starts("<!ENTITY")) {
                    i += 8; ws();
                    size_t b = i;
                    while (i < n && !isspace((unsigned char)s[i])) i++;
                    std::string name(s + b, i - b);
                    ws();
                    if (i >= n || (s[i] != '"' && s[i] != '\'')) return fail("bad ENTITY declaration");
                    char q = s[i++];
                    b = i;
                    while (i < n && s[i] != q) i++;
                    entities[name] = std::string(s + b, i - b);
                    i++;
                    while (i < n && s[i] != '>') i++;
                    i++;
                } else {
                    while (i < n && s[i] != '>') i++;
                    i++;
                }
            }
        }
        ws();
        if (i < n && s[i] == '>') i++;
        return true;
    }
    bool element(XNode &node) {
        if (s[i] != '<') return fail("expected an element");
        i++;
        size_t b = i;
        while (i < n && (isalnum((unsigned char)s[i]) || s[i] == '_' || s[i] == ':' || s[i] == '-' || s[i] == '.')) i++;
        node.name.assign(s + b, i - b);
        for (;;) {
            ws();
            if (i >= n) return fail("unterminated element " + node.name);
            if (starts("/>")) { i += 2; return true; }
            if (s[i] == '>') { i++; break; }
            b = i;
            while (i < n && s[i] != '=' && !isspace((unsigned char)s[i])) i++;
            std::string an(s + b, i - b);
            ws();
            if (s[i] != '=') return fail("attribute without value in " + node.name);
            i++; ws();
            if (s[i] != '"' && s[i] != '\'') return fail("unquoted attribute in " + node.name);
            char q = s[i++];
            b = i;
            while (i < n && s[i] != q) i++;
            node.attrs.emplace_back(an, expand(std::string(s + b, i - b)));
            i++;
        }
        // content
        for (;;) {
            if (i >= n) return fail("unterminated element " + node.name);
            if (starts("</")) {
                i += 2;
                size_t e = i;
                while (e < n && s[e] != '>') e++;
                std::string close(s + i, e - i);
                while (!close.empty() && isspace((unsigned char)close.back())) close.pop_back();
                if (close != node.name) return fail("mismatched </" + close + "> for <" + node.name + ">");
                i = e + 1;
                node.text = expand(node.text);
                return true;
            }
            if (starts("<!--")) {
                size_t e = std::string(s + i).find("-->");
                if (e == std::string::npos) return fail("unterminated comment");
                i += e + 3;
                continue;
            }
            if (starts("<![CDATA[")) {
                size_t e = std::string(s + i).find("]]>");
                if (e == std::string::npos) return fail("unterminated CDATA");
                node.text.append(s + i + 9, e - 9);
                i += e + 3;
                continue;
            }
            if (s[i] == '<') {
                node.kids.emplace_back(new XNode());
                if (!element(*node.kids.back())) return false;
                continue;
            }
            node.text += s[i++];
        }
    }
};

double num(const char *v, double dflt = 0.) { return v ? strtod(v, nullptr) : dflt; }
double textnum(const XNode &x) { return strtod(x.text.c_str(), nullptr); }

// SEC_UNITS attribute scaling (Common/System/UnitsController.cpp:455-465)
double time_units(const XNode &x) {
    const char *u = x.attr("units");
    if (!u) return 1.;
    std::string s(u);
    if (s == "msec" || s == "ms") return 1.e-3;
    if (s == "microsec" || s == "us") return 1.e-6;
    return 1.;
}

// LENGTH_UNITS and VELOCITY_UNITS attribute scaling to mm and mm/s
double length_units(const XNode &x) {
    const char *u = x.attr("units");
    if (!u) return 1.;
    std::string s(u);
    if (s == "m") return 1.e3;
    if (s == "cm") return 10.;
    if (s == "microns") return 1.e-3;
    if (s == "in") return 25.4;
    return 1.;
}
double velocity_units(const XNode &x) {
    const char *u = x.attr("units");
    if (!u) return 1.;
    std::string s(u);
    if (s == "m/sec" || s == "m/s") return 1.e3;
    if (s == "mm/msec" || s == "mm/ms") return 1.e3;
    if (s == "cm/sec" || s == "cm/s") return 10.;
    return 1.;
}

struct Reader {
    MpmProblem &p;
    std::map<std::string, int> matByName;
    explicit Reader(MpmProblem &pp) : p(pp) {}
    bool fail(const std::string &m) { if (p.err.empty()) p.err = m; return false; }

    int mat_id(const XNode &x, const char *numAttr) {
        const char *mn = x.attr("matname");
        if (mn) {
            auto it = matByName.find(mn);
            return it == matByName.end() ? -1 : it->second;
        }
        return (int)num(x.attr(numAttr), 0.);
    }

    bool header(const XNode &h) {
        for (auto &k : h.kids)
            if (k->name == "Analysis") {
                int a = (int)textnum(*k);
                if (a != MPMGPU_PLANE_STRAIN && a != MPMGPU_THREED)
                    return fail("<Analysis> " + std::to_string(a) + " is not on this path (10 plane strain or 12 3D)");
                p.analysis = a;
                p.dim = a == MPMGPU_THREED ? 3 : 2;
            }
        return true;
    }

    bool mpmheader(const XNode &h) {
        for (auto &kp : h.kids) {
            const XNode &k = *kp;
            if (k.name == "MPMMethod") {
                if ((int)textnum(k) != 2) return fail("only MPMMethod 2 (USAVG) is on this path");
            } else if (k.name == "MatlPtsPerElement") {
                int np = (int)textnum(k);
                int side = p.dim == 3 ? (int)(cbrt((double)np) + 0.5) : (int)(sqrt((double)np) + 0.5);
                if ((p.dim == 3 && side * side * side != np) || (p.dim == 2 && side * side != np))
                    return fail("MatlPtsPerElement must be a square (2D) or a cube (3D)");
                p.ptsPerSide = side;
            } else if (k.name == "Timing") {
                double sc = time_units(k);
                if (k.attr("CFL")) p.cfl = num(k.attr("CFL"));
                if (k.attr("step")) p.timestepIn = num(k.attr("step")) * sc;
                if (k.attr("max")) p.maxTime = num(k.attr("max")) * sc;
            } else if (k.name == "TimeFactor") {
                p.cfl = textnum(k);
            } else if (k.name == "MaxTime") {
                p.maxTime = textnum(k) * time_units(k);
            } else if (k.name == "TimeStep") {
                p.timestepIn = textnum(k) * time_units(k);
            } else if (k.name == "GIMP") {
                const char *t = k.attr("type");
                std::string v = t ? t : "uGIMP";
                if (v == "uGIMP" || v == "GIMP" || v == "1") p.run.shape = MPMGPU_SHAPE_UGIMP;
                else if (v == "Dirac" || v == "Classic" || v == "0") p.run.shape = MPMGPU_SHAPE_LINEAR;
                else if (v == "lCPDI" || v == "CPDI" || v == "2") p.run.shape = MPMGPU_SHAPE_LCPDI;
                else return fail("GIMP type " + v + " is not on this path");
            } else if (k.name == "MultiMaterialMode") {
                // MPMReadHandler.cpp:603-650 (Normals) and :896-958,1792-1812 (<Friction>)
                p.mmRequested = true;
                p.normalsGiven = false;
                if (k.attr("Normals")) {
                    double v = num(k.attr("Normals"));
                    if (!(v < 6.5)) return fail("Normals attribute on MultiMaterialMode must be 0 to 6.");
                    // the reference's ranges, in which a value below -0.5 falls into 1
                    int m = v < 0.5 && v > -0.5 ? 0 : v < 1.5 ? 1 : v < 2.5 ? 2 : v < 3.5 ? 3 : v < 4.5 ? 4 : v < 5.5 ? 5 : 6;
                    p.run.normalsMethod = m;
                    p.normalsGiven = true;
                }
                // RigidBias only changes contact with rigid materials, which this path does not have
                if (k.attr("Polar")) p.polarAngle = num(k.attr("Polar"));
                if (k.attr("Azimuth")) p.azimuthAngle = num(k.attr("Azimuth"));
                for (auto &c : k.kids) {
                    if (c->name == "ContactPosition") return fail("<ContactPosition> (contact by positions) is not on this path");
                    if (c->name != "Friction") continue;
                    if (c->attr("lawname")) {
                        p.contactLawName = c->attr("lawname");
                    } else if (c->attr("law")) {
                        p.contactLawNum = (int)num(c->attr("law"));
                    } else {
                        // old style: the text is a coefficient turned into an automatic law
                        // (ContactLaw::ConvertOldStyleToContactLaw, ContactLaw.cpp:168-207)
                        double f = textnum(*c);
                        MaterialInput law;
                        law.name = "automatic contact law";
                        if (f < -10.) law.type = MPMGPU_MAT_CONTACT;
                        else if (f < 0.) { law.type = MPMGPU_MAT_COULOMB; law.friction = -1.; }
                        else if (f < 10.) { law.type = MPMGPU_MAT_COULOMB; law.friction = f; }
                        else return fail("imperfect-interface contact is not on this path");
                        p.autoLaws.push_back(law);
                        p.contactLawNum = -(int)p.autoLaws.size();
                    }
                }
            } else if (k.name == "ArchiveRoot") {
                p.archiveRoot = k.text;
            } else if (k.name == "MPMArchiveOrder") {
                // ArchiveData::SetMPMOrder (ArchiveData.cpp:2131); normalised when an archive is written
                std::string t;
                for (char ch : k.text) if (!std::isspace((unsigned char)ch)) t.push_back(ch);
                p.archiveOrder = t;
            } else if (k.name == "MPMArchive") {
                // <MPMArchive result="..." setting="Y|N"/> (MPMReadHandler.cpp:674-800): one item by
                // name; history1-4 set or clear one bit of the history byte ('0' | bits, 'N' when none)
                static const struct { const char *name; int item; } items[] = {
                    {"velocity", 2}, {"stress", 3}, {"strain", 4}, {"rotstrain", 17}, {"plasticstrain", 5},
                    {"strainenergy", 12}, {"workenergy", 7}, {"plasticenergy", 9}, {"heatenergy", 15},
                    {"lp", 19}, {"wp", 20}, {"temperature", 8}, {"concentration", 14}, {"porepressure", 14},
                    {"elementcrossings", 16}, {"damagenormal", 18}, {"size", 24}};
                const char *r = k.attr("result");
                const char *set = k.attr("setting");
                const char setting = set && set[0] ? set[0] : 'Y';
                if (!r) return fail("<MPMArchive> needs a result attribute");
                std::string name;
                for (const char *c = r; *c; c++) name.push_back((char)std::tolower((unsigned char)*c));
                while ((int)p.archiveOrder.size() < 25) p.archiveOrder.push_back('N');
                if (name.size() == 8 && name.compare(0, 7, "history") == 0 && name[7] >= '1' && name[7] <= '4') {
                    char h = p.archiveOrder[13];
                    if (h == 'Y') h = 0x31;
                    else if (h == 'N') h = 0x30;
                    const char bit = (char)(1 << (name[7] - '1'));
                    if (setting == 'Y') h |= bit;
                    else if (h & bit) h -= bit;
                    p.archiveOrder[13] = h == 0x30 ? 'N' : h;
                } else {
                    int item = -1;
                    for (const auto &it : items) if (name == it.name) item = it.item;
                    if (name.compare(0, 7, "history") == 0) item = 21;   // histories past 4: not carried
                    if (item < 0) return fail(std::string("<MPMArchive> result '") + r + "' is not an archive item here");
                    p.archiveOrder[item] = setting;
                }
            } else if (k.name == "ArchiveTime") {
                p.archiveTime = textnum(k) * time_units(k);
            } else if (k.name == "LeaveLimit") {
                p.leaveLimit = (int)textnum(k);           // MPMReadHandler.cpp:837-843
            } else if (k.name == "DeleteLimit") {
                p.deleteLimit = (int)textnum(k);          // MPMReadHandler.cpp:844-850
            } else if (k.name == "Damping" || k.name == "PDamping") {
                // MPMReadHandler.cpp:311-350: a constant in 1/s (BodyForce damping, pdamping)
                if (k.attr("function")) return fail("<" + k.name + " function=> (damping that varies in time) is not on this path");
                if (k.attr("PIC")) return fail("PIC attribute on damping remove; use PeriodicXPIC task intead");
                (k.name == "Damping" ? p.run.gridDamping : p.run.particleDamping) = textnum(k);
            } else if (k.name == "ExactTractions") {
                return fail("<ExactTractions> is not on this path");
            } else if (k.name == "FeedbackDamping" || k.name == "PFeedbackDamping" || k.name == "XPIC") {
                return fail("<" + k.name + "> is not on this path");
            }
            // archive, global-archive and output-only settings do not change the step
        }
        return true;
    }

    bool mesh(const XNode &m) {
        for (auto &gp : m.kids) {
            if (gp->name != "Grid") continue;
            const XNode &g = *gp;
            double ls = length_units(g);
            double ext[6] = {num(g.attr("xmin")) * ls, num(g.attr("xmax")) * ls, num(g.attr("ymin")) * ls,
                             num(g.attr("ymax")) * ls, num(g.attr("zmin")) * ls, num(g.attr("zmax")) * ls};
            double cell[3] = {-1., -1., -1.};
            int nc[3] = {0, 0, 0};
            if (g.attr("thickness")) p.thickness = num(g.attr("thickness"));
            const char *axes[3] = {"Horiz", "Vert", "Depth"};
            for (auto &ap : g.kids)
                for (int a = 0; a < 3; a++)
                    if (ap->name == axes[a]) {
                        if (ap->attr("cellsize")) cell[a] = num(ap->attr("cellsize")) * ls;
                        if (ap->attr("nx")) nc[a] = (int)num(ap->attr("nx"));
                    }
            p.haveGrid = true;
            p.gxmin = ext[0]; p.gxmax = ext[1]; p.gymin = ext[2]; p.gymax = ext[3]; p.gzmin = ext[4]; p.gzmax = ext[5];
            p.cellx = cell[0]; p.celly = cell[1]; p.cellz = cell[2];
            p.nhoriz = nc[0]; p.nvert = nc[1]; p.ndepth = nc[2];
            if (!mpmsetup::build_grid(p)) return false;
        }
        return true;
    }

    bool body(const XNode &b) {
        int mat = mat_id(b, "mat");
        if (mat <= 0) return fail("Positive material ID needed in <Body> element.");
        double vel[3] = {num(b.attr("vx")), num(b.attr("vy")), num(b.attr("vz"))};
        double thick = b.attr("thick") ? num(b.attr("thick")) : -1.;
        int savedSide = p.ptsPerSide;
        for (auto &sp : b.kids) {
            const XNode &s = *sp;
            if (s.name == "MatlPtsPerElement") {
                int np = (int)textnum(s);
                p.ptsPerSide = p.dim == 3 ? (int)(cbrt((double)np) + 0.5) : (int)(sqrt((double)np) + 0.5);
                continue;
            }
            double ls = length_units(s);
            double ext[6] = {num(s.attr("xmin")) * ls, num(s.attr("xmax")) * ls, num(s.attr("ymin")) * ls,
                             num(s.attr("ymax")) * ls, num(s.attr("zmin")) * ls, num(s.attr("zmax")) * ls};
            int kind = -1, axis = 0;
            if (s.name == "Box") {
                kind = 0;
                if (const char *ax = s.attr("axis")) {
                    std::string a(ax);
                    axis = (a == "x" || a == "X" || a == "1") ? 1 : (a == "y" || a == "Y" || a == "2") ? 2
                         : (a == "z" || a == "Z" || a == "3") ? 3 : 0;
                }
            } else if (s.name == "Cylinder") {
                kind = 0;
                axis = 3;
                if (const char *ax = s.attr("axis")) {
                    std::string a(ax);
                    axis = (a == "x" || a == "X" || a == "1") ? 1 : (a == "y" || a == "Y" || a == "2") ? 2 : 3;
                }
            } else if (s.name == "Rect") kind = 1;
            else if (s.name == "Oval") kind = 2;
            else if (s.name == "Sphere") kind = 3;
            else return fail("<" + s.name + "> in <Body> is not on this path");
            if (!s.kids.empty()) return fail("nested shapes (holes) in <Body> are not on this path");
            if (!mpmsetup::add_body(p, mat, kind, ext, axis, vel, thick)) return false;
        }
        p.ptsPerSide = savedSide;
        return true;
    }

    bool points(const XNode &mp) {
        // <PointList><mp matl= lpx=><pt x= y= z=/><vel x= y=/><mass m=/></mp></PointList>
        for (auto &xp : mp.kids) {
            const XNode &x = *xp;
            if (x.name == "Body") { if (!body(x)) return false; continue; }
            if (x.name == "PointList") { if (!points(x)) return false; continue; }
            if (x.name != "mp") return fail("<" + x.name + "> in <MaterialPoints> is not on this path");
            int mat = x.attr("matname") ? mat_id(x, "matl") : (int)num(x.attr("matl") ? x.attr("matl") : x.attr("mat"));
            double lp[3] = {num(x.attr("lpx"), 0.5), num(x.attr("lpy"), 0.5), num(x.attr("lpz"), 0.5)};
            double thick = x.attr("thick") ? num(x.attr("thick")) : -1.;
            double pos[3] = {0., 0., 0.}, vel[3] = {0., 0., 0.}, mass = -1.;
            for (auto &cp : x.kids) {
                const XNode &c = *cp;
                if (c.name == "pt") {
                    double sc = length_units(c);
                    pos[0] = num(c.attr("x")) * sc; pos[1] = num(c.attr("y")) * sc; pos[2] = num(c.attr("z")) * sc;
                } else if (c.name == "vel") {
                    double sc = velocity_units(c);
                    vel[0] = num(c.attr("x")) * sc; vel[1] = num(c.attr("y")) * sc; vel[2] = num(c.attr("z")) * sc;
                } else if (c.name == "mass") {
                    double m = num(c.attr("m"), -1.);
                    if (m > 0.) mass = m;
                }
            }
            if (!mpmsetup::add_point(p, mat, pos, vel, lp, thick, mass)) return false;
        }
        return true;
    }

    bool material(const XNode &m) {
        MaterialInput in;
        in.type = (int)num(m.attr("Type"));
        in.name = m.attr("Name") ? m.attr("Name") : "";
        for (auto &kp : m.kids) {
            const XNode &k = *kp;
            double v = textnum(k);
            const std::string &n = k.name;
            if (n == "rho") in.rho = v;
            else if (n == "E" || n == "Etens") { in.E = v; in.readE = true; }
            else if (n == "nu") { in.nu = v; in.readNu = true; }
            else if (n == "G" || n == "G1") { in.G = v; in.readG = true; }
            else if (n == "K" || n == "Kbulk") in.K = v;
            else if (n == "Lame") in.Lame = v;
            else if (n == "G2") in.G2 = v;
            else if (n == "yield") in.yield = v;
            else if (n == "yieldMin") in.yieldMin = v;
            else if (n == "Ep") in.Ep = v;
            else if (n == "Khard") in.Khard = v;
            else if (n == "coeff") in.friction = v;
            else if (n == "coeffStatic") in.frictionStatic = v;
            else if (n == "UJOption") in.ujOption = (int)v;
            else if (n == "PDamping") {
                // MPMReadHandler.cpp:311-343: a constant in 1/s; no function of time for a material
                if (k.attr("function")) return fail("Material damping function of time is not allowed");
                if (k.attr("PIC")) return fail("PIC attribute on damping remove; use PeriodicXPIC task intead");
                in.pdamping = v;
            }
            else if (n == "Friction" || n == "SharedMaterialField" || n == "ArtificialVisc" || n == "ArtificialViscosity")
                return fail("<" + n + "> in <Material> is not on this path");
            else if (n == "Hardening") {
                if (k.text.find("Linear") == std::string::npos && (int)v != 1 && !k.text.empty() && k.text.find_first_not_of(" \t\r\n") != std::string::npos)
                    return fail("only linear hardening is on this path");
            }
            // thermal, transport and output-only properties do not act on this path
        }
        int id = (int)p.matIn.size() + 1;
        if (!in.name.empty()) matByName[in.name] = id;
        mpmsetup::add_material(p, in);
        return true;
    }

    // CommonReadHandler::ReadGridPoint (Common/Read_XML/CommonReadHandler.cpp:611-658): a number
    // times the length scale, or the mesh's min or max on the axis plus an optional number of
    // minimum cell sizes ("max", "max+", "max-", "max3", "min+0.5", ...)
    static double grid_point(const char *v, double ls, double amin, double amax, double minCell) {
        const size_t len = strlen(v);
        if (len >= 3 && v[0] == 'm' && ((v[1] == 'a' && v[2] == 'x') || (v[1] == 'i' && v[2] == 'n'))) {
            double d = 0.;
            if (len == 4) {
                if (v[3] == '-') d = -1.;
                else if (v[3] == '+') d = 1.;
                else if (v[3] >= '0' && v[3] <= '9') d = (float)(v[3] - '0');
            } else if (len > 4) {
                d = strtod(v + 3, nullptr);
            }
            return (v[1] == 'a' ? amax : amin) + d * minCell;
        }
        return strtod(v, nullptr) * ls;
    }

    // DbleEqual (Common/System/CommonUtilities.cpp:46-61)
    static bool dble_equal(double a, double b) {
        double diff = std::fabs(a - b);
        if (diff <= 1.e-16) return true;
        double l = std::max(std::fabs(a), std::fabs(b));
        return diff <= l * 1.e-7;
    }

    // a <BCLine> (2D) or <BCBox> (3D) of <GridBCs> or <ParticleBCs> (Read_MPM/Generators.cpp:831-891)
    struct BcShape {
        bool line = false;
        double lo[3] = {0., 0., 0.}, hi[3] = {0., 0., 0.};
        double dtol = 0., distSq = 0.;
        // LineController::ContainsPoint (LineController.cpp:49-66); BoxController::ContainsPoint,
        // axis 0 (BoxController.cpp:70-74)
        bool contains(double x, double y, double z) const {
            if (!line)
                return x <= hi[0] && x >= lo[0] && y <= hi[1] && y >= lo[1] && z <= hi[2] && z >= lo[2];
            double dx = hi[0] - lo[0], dy = hi[1] - lo[1];
            double px = x - lo[0], py = y - lo[1];
            double ypr = -dy * px + dx * py;
            if (ypr > dtol || ypr < -dtol) return false;
            double xpr = dx * px + dy * py;
            return !(xpr < 0. || xpr > distSq);
        }
    };

    // ElementBase::gridTolerance, 0.1 of the shortest element side (ElementBase.cpp:60-84,
    // ElementBase3D.cpp:30-44); GetMinimumCellSize divides it by 0.1 again (:150)
    double grid_tolerance() const {
        const MpmGridDesc &g = p.grid;
        double gtol = 1.e20;
        const long nn[3] = {g.nx + 1L, g.ny + 1L, p.dim == 3 ? g.nz + 1L : 1L};
        const double o[3] = {g.xmin, g.ymin, g.zmin}, d[3] = {g.dx, g.dy, g.dz};
        for (int c = 0; c < p.dim; c++)
            for (long i = 0; i + 1 < nn[c]; i++)
                gtol = std::min(gtol, 0.1 * ((o[c] + (double)(i + 1) * d[c]) - (o[c] + (double)i * d[c])));
        return gtol;
    }

    bool bc_shape(const XNode &s, BcShape &sh) {
        const int dim = p.dim;
        const bool line = s.name == "BCLine", box = s.name == "BCBox";
        if (!line && !box) return fail("<" + s.name + "> in <GridBCs> or <ParticleBCs> is not on this path");
        if (line && dim != 2) return fail("<BCLine> is for 2D analyses");
        if (box && dim != 3) return fail("<BCBox> is for 3D analyses");
        const double gtol = grid_tolerance(), minCell = gtol / 0.1;
        const double ls = length_units(s);
        sh.line = line;
        double tol = gtol;
        // ShapeController::SetProperty (ShapeController.cpp:105-125): x1/xmin ... z2/zmax,
        // the later attribute of a pair wins
        for (auto &a : s.attrs) {
            const std::string &k = a.first;
            const char *v = a.second.c_str();
            for (int c = 0; c < 3; c++) {
                const char ax = "xyz"[c];
                const double amin = p.meshLim[2 * c], amax = p.meshLim[2 * c + 1];
                if (k == std::string(1, ax) + "1" || k == std::string(1, ax) + "min")
                    sh.lo[c] = grid_point(v, ls, amin, amax, minCell);
                else if (k == std::string(1, ax) + "2" || k == std::string(1, ax) + "max")
                    sh.hi[c] = grid_point(v, ls, amin, amax, minCell);
            }
            if (k == "tolerance" && line)      // LineController::SetProperty (LineController.cpp:91-105)
                tol = v[0] == '*' ? strtod(v + 1, nullptr) * minCell : strtod(v, nullptr) * ls;
            if (k == "axis" && box && std::string(v) != "0") {     // BoxController::SetProperty (:22-36)
                const std::string av(v);
                if (av == "x" || av == "X" || av == "1" || av == "y" || av == "Y" || av == "2" || av == "z" ||
                    av == "Z" || av == "3")
                    return fail("<BCBox axis=> (cylinder node sets) is not on this path");
                return fail("Box axis must be x, y, z, 1, 2, or 3");
            }
        }
        if (line) {
            // LineController::FinishSetup (LineController.cpp:33-42)
            sh.distSq = (sh.hi[0] - sh.lo[0]) * (sh.hi[0] - sh.lo[0]) + (sh.hi[1] - sh.lo[1]) * (sh.hi[1] - sh.lo[1]);
            sh.dtol = std::sqrt(sh.distSq) * tol;
        } else {
            // ShapeController::FinishSetup (ShapeController.cpp:164-197): ordered, nonzero extents
            for (int c = 0; c < 3; c++) {
                if (sh.lo[c] > sh.hi[c]) std::swap(sh.lo[c], sh.hi[c]);
                if (dble_equal(sh.lo[c], sh.hi[c]))
                    return fail(std::string("Box: ") + "xyz"[c] + "max cannot equal " + "xyz"[c] + "min in input parameters.");
            }
        }
        return true;
    }

    // common attributes of <DisBC> and <LoadBC>: the style, and ftime converted as
    // BoundaryCondition::SetBCFirstTime does (BoundaryCondition.cpp:200-206): ms -> s, or
    // 1/ms -> 1/s for the sine and cosine styles
    static double first_time(int style, double ftime) {
        return (style == 3 || style == 4) ? ftime * 1.e3 : ftime * 1.e-3;
    }

    // <GridBCs>: <BCLine> (2D) or <BCBox> (3D) node sets holding <DisBC> velocity conditions
    // (MPMReadHandler, Read_MPM/Generators.cpp:831-1010). Each DisBC adds one condition per node
    // the shape contains, nodes in node order (ShapeController::nextNode,
    // Common/Read_XML/ShapeController.cpp:231-244), so the list runs DisBC by DisBC.
    bool gridbcs(const XNode &gb) {
        if (!p.gridBuilt) return fail("<GridBCs> needs the <Mesh> first");
        const MpmGridDesc &g = p.grid;
        const int dim = p.dim;
        const long nnx = g.nx + 1, nny = g.ny + 1, nnz = dim == 3 ? g.nz + 1 : 1;
        for (auto &sp : gb.kids) {
            const XNode &s = *sp;
            BcShape sh;
            if (!bc_shape(s, sh)) return false;
            for (auto &bp : s.kids) {
                const XNode &b = *bp;
                if (b.name != "DisBC")
                    return fail("<" + b.name + "> in <" + s.name + "> is not on this path");
                int dof = 0, style = 1, id = 0;
                double vel = 0., ftime = 0., angle = 0.;
                for (auto &a : b.attrs) {
                    const std::string &k = a.first;
                    const char *v = a.second.c_str();
                    if (k == "dir") dof = (int)strtol(v, nullptr, 10);
                    else if (k == "disp" || k == "vel") vel = strtod(v, nullptr);
                    else if (k == "time" || k == "ftime") ftime = strtod(v, nullptr);
                    else if (k == "style") style = (int)strtol(v, nullptr, 10);
                    else if (k == "angle") angle = strtod(v, nullptr);
                    else if (k == "id") id = (int)strtol(v, nullptr, 10);
                    else if (k == "function" || k == "gradfxn" || k == "dispfxn" || k == "depth" || k == "angle2")
                        return fail("<DisBC " + k + "=> is not on this path");
                }
                if (dim == 3 && dof != 1 && dof != 2 && dof != 3 && dof != 12 && dof != 13 && dof != 23 && dof != 123)
                    return fail("'dir' in DisBC element must be 1, 2, 3, 12, 13, 23, or 123 for 3D analyses.");
                if (dim == 2 && dof != 1 && dof != 2 && dof != 12)
                    return fail("'dir' in DisBC element must be 1, 2, or 12 for 2D analyses.");
                if (id > 0) return fail("'id' for velocity boundary conditions must be <= 0.");
                if (dof > 10) {
                    if (!dble_equal(angle, 0.)) return fail("skewed velocity conditions are not on this path");
                    dof = (dof == 12 || dof == 13) ? 1 : (dof == 23 ? 2 : 3);
                }
                if (style < 1 || style > 4) return fail("velocity condition styles other than 1 to 4 are not on this path");
                // BoundaryCondition::SetBCValue (BoundaryCondition.cpp:182-188): linear values per ms to per s
                MpmVelocityBC bc{};
                bc.dir = dof;
                bc.style = style;
                bc.value = style == 2 ? vel * 1.e3 : vel;
                bc.ftime = first_time(style, ftime);
                for (long k = 0; k < nnz; k++)
                    for (long j = 0; j < nny; j++)
                        for (long i = 0; i < nnx; i++)
                            if (sh.contains(g.xmin + (double)i * g.dx, g.ymin + (double)j * g.dy,
                                            dim == 3 ? g.zmin + (double)k * g.dz : 0.)) {
                                bc.node = (k * nny + j) * nnx + i;
                                p.velBCs.push_back(bc);
                            }
            }
        }
        return true;
    }

    // <ParticleBCs>: <LoadBC> external loads and <TractionBC> tractions on the particles a <BCLine> (2D) or <BCBox> (3D)
    // contains, particles in mpm[] order (ShapeController::nextParticle, ShapeController.cpp:277-290);
    // after <net/> the load is divided by the shape's particle count, <perParticle/> returns to
    // one load per particle (Generators.cpp:1012-1023, ShapeController::setNetBC, :292-309)
    bool particlebcs(const XNode &pb) {
        const int dim = p.dim;
        const int64_t n = p.nparticles();
        for (auto &sp : pb.kids) {
            const XNode &s = *sp;
            BcShape sh;
            if (!bc_shape(s, sh)) return false;
            std::vector<int64_t> inside;
            for (int64_t i = 0; i < n; i++)
                if (sh.contains(p.d[MPF_POS_X][i], p.d[MPF_POS_Y][i], dim == 3 ? p.d[MPF_POS_Z][i] : 0.))
                    inside.push_back(i);
            bool net = false;
            for (auto &bp : s.kids) {
                const XNode &b = *bp;
                if (b.name == "net") { net = true; continue; }
                if (b.name == "perParticle") { net = false; continue; }
                if (b.name != "LoadBC" && b.name != "TractionBC")
                    return fail("<" + b.name + "> in <" + s.name + "> is not on this path");
                int dof = 0, style = 1, face = 1;
                double load = 0., ftime = 0.;
                for (auto &a : b.attrs) {
                    const std::string &k = a.first;
                    const char *v = a.second.c_str();
                    if (k == "dir") dof = (int)strtol(v, nullptr, 10);
                    else if (k == "style") style = (int)strtol(v, nullptr, 10);
                    else if (k == "face") face = (int)strtol(v, nullptr, 10);
                    else if (k == "load" || k == "value" || k == "stress") load = strtod(v, nullptr);
                    else if (k == "ftime" || k == "time" || k == "bath") ftime = strtod(v, nullptr);
                    else if (k == "function") return fail("<" + b.name + " function=> is not on this path");
                }
                if (b.name == "TractionBC") {
                    // Generators.cpp:1139-1167,1189-1202: checks, then one condition per particle
                    // (not divided by the particle count); the value in MPa becomes Pa
                    // (MatPtLoadBC::SetBCValue, MatPtLoadBC.cpp:202-206)
                    if (dim == 3 && !((dof >= 1 && dof <= 3) || dof == 11))
                        return fail("'dir' in <TractionBC> must be 1-3 or 11 for 3D analyses.");
                    if (dim == 3 && (face < 1 || face > 6))
                        return fail("'face' in <TractionBC>, <ConcFluxBC>, or <HeatFluxBC> element must be 1 to 6 for 3D analyses.");
                    if (dim == 2 && !(dof == 1 || dof == 2 || dof == 11 || dof == 12))
                        return fail("'dir' in <LoadBC> or <ConcFluxBC> must be 1 or 2 and in <TractionBC> must be 1, 2, 11, or 12 for 2D analyses.");
                    if (dim == 2 && (face < 1 || face > 4))
                        return fail("'face' in <TractionBC>, <ConcFluxBC>, or <HeatFluxBC> element must be 1 to 4 for 2D analyses.");
                    if (style == 5) return fail("Tractions boundary conditions cannot use silent style.");
                    if (style < 1 || style > 4) return fail("traction styles other than 1 to 4 are not on this path");
                    double v = 1.e6 * load;
                    if (style == 2) v = v * 1.e3;
                    MpmParticleTraction tr{};
                    tr.dir = dof;
                    tr.face = face;
                    tr.style = style;
                    tr.value = v;
                    tr.ftime = first_time(style, ftime);
                    for (int64_t i : inside) {
                        tr.particle = i;
                        tr.thickness = dim == 2 ? p.pthick[(size_t)i] : 1.;
                        p.tractions.push_back(tr);
                    }
                    continue;
                }
                if (dim == 3 && (dof < 1 || dof > 3)) return fail("'dir' in <LoadBC> must be 1-3 for 3D analyses.");
                if (dim == 2 && (dof < 1 || dof > 2))
                    return fail("'dir' in <LoadBC> or <ConcFluxBC> must be 1 or 2 and in <TractionBC> must be 1, 2, 11, or 12 for 2D analyses.");
                if (style < 1 || style > 4) return fail("load styles other than 1 to 4 are not on this path");
                // MatPtLoadBC::SetBCValue (MatPtLoadBC.cpp:202-206): N -> uN, then the linear
                // style's per ms -> per s; a net load is shared by the shape's particles
                const double normalize = net && !inside.empty() ? (double)inside.size() : 1.;
                double v = 1.e6 * (load / normalize);
                if (style == 2) v = v * 1.e3;
                MpmParticleLoad ld{};
                ld.dir = dof;
                ld.style = style;
                ld.value = v;
                ld.ftime = first_time(style, ftime);
                for (int64_t i : inside) {
                    ld.particle = i;
                    p.loads.push_back(ld);
                }
            }
        }
        return true;
    }

    // <CustomTasks>: <Schedule name="PeriodicXPIC"> with <Parameter name="...">value</Parameter>
    // (MPMReadHandler.cpp:1682-1690, SetCustomTasks.cpp:27-160, PeriodicXPIC::InputParam
    // PeriodicXPIC.cpp:60-157: periodicTime in ms); other tasks are not on this path
    bool customtasks(const XNode &ct) {
        for (auto &sp : ct.kids) {
            const XNode &sc = *sp;
            if (sc.name != "Schedule") continue;
            const char *nm = sc.attr("name");
            const std::string task = nm ? nm : "";
            if (task != "PeriodicXPIC") return fail("custom task " + task + " is not on this path");
            if (p.xpicTask) return fail("more than one PeriodicXPIC task is not on this path");
            p.xpicTask = true;
            for (auto &pp : sc.kids) {
                const XNode &pr = *pp;
                if (pr.name != "Parameter") continue;
                const char *an = pr.attr("name");
                if (!an) return fail("Custom task parameter has no name.");
                std::string key = an;
                if (!key.empty() && key.back() >= '0' && key.back() <= '9') key.pop_back();   // parameter number
                const char *v = pr.text.c_str();
                if (key == "XPICOrder") p.xpicOrderIn = (int)strtol(v, nullptr, 10);
                else if (key == "FMPMOrder") p.fmpmOrderIn = (int)strtol(v, nullptr, 10);
                else if (key == "periodicSteps") p.xpicStepsIn = (int)strtol(v, nullptr, 10);
                else if (key == "periodicTime") p.xpicTimeIn = strtod(v, nullptr) * 1.e-3;
                else if (key == "periodicCFL") p.xpicCFLIn = strtod(v, nullptr);
                else if (key == "verbose") continue;
                else if (key == "DynamicConverge") {
                    if (strtod(v, nullptr) > 0.) return fail("PeriodicXPIC DynamicConverge is not on this path");
                } else if (key == "periodicTimeConduction" || key == "periodicCFLConduction" ||
                           key == "periodicStepsConduction" || key == "periodicTimeDiffusion" ||
                           key == "periodicCFLDiffusion" || key == "periodicStepsDiffusion")
                    return fail("PeriodicXPIC for conduction and diffusion is not on this path");
                else return fail("Unrecognized custom task parameter: '" + std::string(an) + "'");
            }
        }
        return true;
    }

    bool run(const XNode &root) {
        if (root.name != "JANFEAInput") return fail("root element must be <JANFEAInput>");
        p.run.shape = MPMGPU_SHAPE_LINEAR;       // ElementBase::useGimp default until <GIMP>
        p.run.mpmMethod = 2;
        p.run.contactLaw = -1;
        // headers and materials first (materials may follow the points in the file)
        for (auto &k : root.kids) if (k->name == "Header" && !header(*k)) return false;
        for (auto &k : root.kids) if (k->name == "MPMHeader" && !mpmheader(*k)) return false;
        for (auto &k : root.kids) if (k->name == "Material" && !material(*k)) return false;
        for (auto &k : root.kids) if (k->name == "Mesh" && !mesh(*k)) return false;
        for (auto &k : root.kids) {
            if (k->name == "MaterialPoints" && !points(*k)) return false;
            if (k->name == "PointList" && !points(*k)) return false;
        }
        for (auto &k : root.kids) {
            const std::string &n = k->name;
            if (n == "GridBCs") {
                if (!gridbcs(*k)) return false;
                continue;
            }
            if (n == "ParticleBCs") {
                if (!particlebcs(*k)) return false;
                continue;
            }
            if (n == "Gravity") {
                // MPMReadHandler.cpp:1627-1660: constant body force components (mm/s^2)
                for (auto &c : k->kids) {
                    const std::string &cn = c->name;
                    if (cn == "BodyXForce") p.run.gravity[0] = textnum(*c);
                    else if (cn == "BodyYForce") p.run.gravity[1] = textnum(*c);
                    else if (cn == "BodyZForce") p.run.gravity[2] = textnum(*c);
                    else return fail("<" + cn + "> in <Gravity> (grid body force functions) is not on this path");
                }
                continue;
            }
            if (n == "CustomTasks") {
                if (!customtasks(*k)) return false;
                continue;
            }
            if (n == "CrackList" || n == "Thermal")
                if (!k->kids.empty() || !k->text.empty())
                    return fail("<" + n + "> is not on this path");
        }
        return mpmsetup::finalize(p);
    }
};

}  // namespace

bool mpm_read_xml(const char *text, MpmProblem &p) {
    Parser ps(text);
    if (!ps.skip_misc()) { p.err = ps.err; return false; }
    XNode root;
    if (!ps.element(root)) { p.err = ps.err; return false; }
    Reader r(p);
    return r.run(root);
}
