// settings.cpp -- parameters.xml reader/writer and derived constants of the timestep.
//
// Reads the same keys, with the same parsing and clamping, as Settings::initializeParameters
// (src/Settings.cpp:84-289); writes the restart parameter file of
// Settings::generateXMLParametersNode (src/Settings.cpp:363-400).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "ggx.h"

namespace {

// Returns the text of the first child element <key>...</key> of the root element, or false.
// The reference uses rapidxml's first_node(name) on the root (src/Settings.cpp:111-283).
bool child_value(const std::string& doc, size_t body_begin, size_t body_end, const char* key,
                 std::string& out) {
    std::string open = std::string("<") + key;
    size_t pos = body_begin;
    int depth = 0;
    while (pos < body_end) {
        size_t lt = doc.find('<', pos);
        if (lt == std::string::npos || lt >= body_end) return false;
        if (doc.compare(lt, 4, "<!--") == 0) {
            size_t e = doc.find("-->", lt + 4);
            if (e == std::string::npos) return false;
            pos = e + 3;
            continue;
        }
        if (doc.compare(lt, 2, "<?") == 0) {
            size_t e = doc.find("?>", lt + 2);
            if (e == std::string::npos) return false;
            pos = e + 2;
            continue;
        }
        size_t gt = doc.find('>', lt);
        if (gt == std::string::npos) return false;
        bool closing = doc[lt + 1] == '/';
        bool selfclose = doc[gt - 1] == '/';
        if (closing) {
            depth--;
            pos = gt + 1;
            continue;
        }
        // element name
        size_t ns = lt + 1, ne = ns;
        while (ne < gt && !isspace((unsigned char)doc[ne]) && doc[ne] != '/' && doc[ne] != '>') ne++;
        std::string name = doc.substr(ns, ne - ns);
        if (depth == 0 && name == key) {
            if (selfclose) { out.clear(); return true; }
            size_t close = doc.find(std::string("</") + key, gt + 1);
            if (close == std::string::npos) return false;
            out = doc.substr(gt + 1, close - gt - 1);
            return true;
        }
        if (!selfclose) depth++;
        pos = gt + 1;
    }
    (void)open;
    return false;
}

unsigned long to_ul(const std::string& s) { return std::stoul(s); }
long to_l(const std::string& s) { return std::stol(s); }
double to_d(const std::string& s) { return std::stod(s); }

}  // namespace

extern "C" int ggx_settings_default(ggx_settings* s) {
    if (!s) return GGX_E_ARG;
    std::memset(s, 0, sizeof(*s));
    // static initialisers, src/Settings.cpp:29-80
    s->grid_coarsement_gradient = 0.95;
    s->convolution_mode = 3;          // E_INVALID_VALUE until read
    s->microstructure_gen_mode = 5;   // E_INVALID_VAL until read
    s->research_project = 3;          // E_DEFAULT
    s->hagb_energy = 1.0;
    s->hagb_mobility = 1.0;
    s->gaussian_kernel_time_slope = 0.8359;
    return GGX_OK;
}

extern "C" int ggx_settings_read_xml(const char* path, ggx_settings* s) {
    if (!path || !s) return GGX_E_ARG;
    std::ifstream f(path);
    if (f.fail()) return GGX_E_IO;  // the reference prints and exit(2)s (Settings.cpp:88-92)
    std::stringstream ss;
    ss << f.rdbuf();
    std::string doc = ss.str();
    ggx_settings_default(s);
    // root element must be <Parameters> (Settings.cpp:104-108)
    size_t p = 0;
    size_t root = std::string::npos;
    while (true) {
        size_t lt = doc.find('<', p);
        if (lt == std::string::npos) return GGX_E_CONFIG;
        if (doc.compare(lt, 2, "<?") == 0) { p = doc.find("?>", lt) + 2; continue; }
        if (doc.compare(lt, 4, "<!--") == 0) { p = doc.find("-->", lt) + 3; continue; }
        root = lt;
        break;
    }
    if (doc.compare(root, 11, "<Parameters") != 0) return GGX_E_CONFIG;
    size_t body_begin = doc.find('>', root) + 1;
    size_t body_end = doc.rfind("</Parameters");
    if (body_end == std::string::npos || body_end < body_begin) return GGX_E_CONFIG;

    std::string v;
    auto has = [&](const char* k) { return child_value(doc, body_begin, body_end, k, v); };
    try {
        if (has("StartTime")) s->start_time = (int32_t)to_ul(v);
        if (has("NumberOfParticles")) s->number_of_particles = (int32_t)to_ul(v);
        if (has("NumberOfPointsPerGrain")) s->number_of_points_per_grain = (int32_t)to_ul(v);
        s->user_def_number_of_points_per_grain = s->number_of_points_per_grain;  // Settings.cpp:122
        if (has("AnalysisTimestep")) s->analysis_timestep = (int32_t)to_ul(v);
        if (has("PlotInterval")) s->plot_interval = (int32_t)to_ul(v);
        if (has("NumberOfTimesteps")) s->number_of_timesteps = (int32_t)to_ul(v);
        if (has("BreakupNumber")) s->breakup_number = (int32_t)to_ul(v);
        if (has("DiscreteSamplingRate")) s->discrete_sampling_rate = (int32_t)to_ul(v);
        if (has("DomainBorderSize")) s->domain_border_size = (int32_t)to_ul(v);
        if (has("MicrostructureGenMode")) {
            long m = to_l(v);
            s->microstructure_gen_mode = (int32_t)(m >= 5 ? 5 : m);
        }
        if (has("ReadFromFilename")) std::snprintf(s->read_from_filename, sizeof(s->read_from_filename), "%s", v.c_str());
        if (has("AdditionalFilename")) std::snprintf(s->additional_filename, sizeof(s->additional_filename), "%s", v.c_str());
        if (has("LatticeType")) {
            long m = to_l(v);
            s->lattice_type = (int32_t)(m >= 2 ? 2 : m);
        }
        if (has("HAGB_Energy")) s->hagb_energy = to_d(v);
        if (has("HAGB_Mobility")) s->hagb_mobility = to_d(v);
        if (has("DislocEnPerM")) s->disloc_en_per_m = to_d(v);
        if (has("Physical_Domain_Size")) s->physical_domain_size = to_d(v);
        if (has("TriplePointDrag")) s->triple_point_drag = to_d(v);
        if (has("UseMobilityModel")) s->use_mobility_model = (int32_t)to_ul(v);
        if (has("IdentifyTwins")) s->identify_twins = to_ul(v) != 0;
        if (has("UseMagneticField")) s->use_magnetic_field = to_ul(v) != 0;
        if (has("IsIsotropicNetwork")) s->is_isotropic_network = to_ul(v) != 0;
        if (has("UseTexture")) s->use_texture = to_ul(v) != 0;
        if (has("MaxMisOrientation")) s->max_misorientation = (double)to_ul(v);
        if (has("ExecuteInParallel")) s->execute_in_parallel = to_ul(v) != 0;
        if (has("MaximumNumberOfThreads")) s->maximum_number_of_threads = (int32_t)to_ul(v);
        if (has("GridCoarsement")) s->grid_coarsement = to_ul(v) != 0;
        if (has("GridCoarsementGradient")) s->grid_coarsement_gradient = to_d(v);
        if (has("ConvolutionMode")) {
            long m = to_l(v);
            s->convolution_mode = (int32_t)(m >= 3 ? 3 : m);
        }
        // ResearchMode is a static bool in the reference (Settings.h:98)
        if (has("ResearchMode")) s->research_mode = to_ul(v) != 0;
        if (has("ResearchProject")) {
            long m = to_l(v);
            s->research_project = (int32_t)(m >= 3 ? 3 : m);
        }
        if (has("ConstantSectorRadius")) s->constant_sector_radius = to_d(v);
        if (has("InterpolatingSectorRadius")) s->interpolating_sector_radius = to_d(v);
        if (has("NeighbourTracking")) s->neighbour_tracking = (int32_t)to_ul(v);
        if (has("UseStoredElasticEnergy")) s->use_stored_elastic_energy = to_ul(v) != 0;
        if (has("GrainScheduler")) {
            long m = to_l(v);
            s->grain_scheduler = (int32_t)(m >= 2 ? 2 : m);
        }
        if (has("DecoupleGrains")) s->decouple_grains = to_ul(v) != 0;
        if (has("GaussianKernelUserDefTimeSlope")) s->gaussian_kernel_time_slope = to_d(v);
        if (has("BoxDefaultStoredElasticEnergy")) s->box_default_stored_elastic_energy = to_d(v);
        if (has("UserDefNumberOfPointsPerGrain")) s->user_def_number_of_points_per_grain = (int32_t)to_ul(v);
    } catch (...) {
        return GGX_E_CONFIG;  // std::stoul/stod would throw in the reference as well
    }
    // strip surrounding whitespace from the file names (rapidxml keeps it)
    auto trim = [](char* c) {
        std::string t(c);
        size_t a = t.find_first_not_of(" \t\r\n"), b = t.find_last_not_of(" \t\r\n");
        t = (a == std::string::npos) ? std::string() : t.substr(a, b - a + 1);
        std::snprintf(c, 512, "%s", t.c_str());
    };
    trim(s->read_from_filename);
    trim(s->additional_filename);
    return GGX_OK;
}

extern "C" int ggx_settings_write_xml(const char* path, const ggx_settings* s, const char* network_file,
                                      int32_t loop, int32_t grains) {
    if (!path || !s) return GGX_E_ARG;
    FILE* f = std::fopen(path, "w");
    if (!f) return GGX_E_IO;
    std::fprintf(f, "<?xml version=\"1.0\" encoding=\"utf-8\"?>\n<Parameters>\n");
    auto I = [&](const char* k, long v) { std::fprintf(f, "\t<%s>%ld</%s>\n", k, v, k); };
    auto D = [&](const char* k, double v) { std::fprintf(f, "\t<%s>%.17g</%s>\n", k, v, k); };
    I("StartTime", loop);
    I("NumberOfParticles", grains);
    I("NumberOfPointsPerGrain", s->number_of_points_per_grain);
    I("AnalysisTimestep", s->analysis_timestep);
    I("PlotInterval", s->plot_interval);
    I("NumberOfTimesteps", s->number_of_timesteps);
    I("BreakupNumber", s->breakup_number);
    I("DiscreteSamplingRate", s->discrete_sampling_rate);
    I("DomainBorderSize", s->domain_border_size);
    I("MicrostructureGenMode", 0);
    std::fprintf(f, "\t<ReadFromFilename>%s</ReadFromFilename>\n", network_file ? network_file : "");
    D("HAGB_Energy", s->hagb_energy);
    D("HAGB_Mobility", s->hagb_mobility);
    D("Physical_Domain_Size", s->physical_domain_size);
    D("TriplePointDrag", s->triple_point_drag);
    I("UseMobilityModel", s->use_mobility_model);
    I("IdentifyTwins", s->identify_twins);
    I("IsIsotropicNetwork", s->is_isotropic_network);
    I("UseTexture", s->use_texture);
    I("ExecuteInParallel", s->execute_in_parallel);
    I("MaximumNumberOfThreads", s->maximum_number_of_threads);
    I("GridCoarsement", s->grid_coarsement);
    D("GridCoarsementGradient", s->grid_coarsement_gradient);
    I("ConvolutionMode", s->convolution_mode);
    I("ResearchMode", s->research_mode);
    I("ResearchProject", s->research_project);
    D("ConstantSectorRadius", s->constant_sector_radius);
    D("InterpolatingSectorRadius", s->interpolating_sector_radius);
    I("NeighbourTracking", s->neighbour_tracking);
    I("UseStoredElasticEnergy", s->use_stored_elastic_energy);
    I("GrainScheduler", s->grain_scheduler);
    std::fprintf(f, "</Parameters>\n");
    std::fclose(f);
    return GGX_OK;
}

// grainhdl::setSimulationParameter (src/grainhdl.cpp:56-136) and find_correctTimestepSize
// (250-274) for MicrostructureGenMode 0.  m_Energy_deltaMAX is never assigned without a
// magnetic field (grainhdl.h:53); it is taken as 0, so the correction factor is 1.
extern "C" int ggx_settings_derive(const ggx_settings* s, ggx_derived* d) {
    if (!s || !d) return GGX_E_ARG;
    if (s->number_of_particles <= 0 || s->number_of_points_per_grain <= 0) return GGX_E_CONFIG;
    if (s->convolution_mode < 0 || s->convolution_mode > 2) return GGX_E_CONFIG;
    std::memset(d, 0, sizeof(*d));
    int R = (int)(std::sqrt((double)s->number_of_particles) * (double)s->number_of_points_per_grain);
    d->real_domain_size = R;
    d->ds = (s->constant_sector_radius * 2 + s->interpolating_sector_radius * 2) / double(R);
    double ppg = (double)s->number_of_points_per_grain;
    if (s->decouple_grains == 1) ppg = (double)s->user_def_number_of_points_per_grain;
    // every convolution mode uses the same formula (grainhdl.cpp:89-123)
    d->dt = 0.8 / double(R * R) * ppg / 2 / 5.0;
    d->time_slope = s->gaussian_kernel_time_slope;
    d->h = 1.0 / double(R);
    d->delta = s->domain_border_size * 1 / double(R);
    d->grid_blowup = s->domain_border_size;
    d->ngridpoints = R + 2 * s->domain_border_size;
    double vmax = 0.0 * d->dt;
    double corr = 0.5 / (vmax / d->h);
    if (corr > 1.0) corr = 1.0;
    d->dt *= corr;
    return GGX_OK;
}
