// GpuSimMediator.h -- the reference-side binding of libvedo_hip.so.
//
// A maintainer adds this header to VEDO (e.g. src/vedo/gpu/) and a `gpu`
// choice next to `nbs`/`leap` in Anne.cpp (Anne.cpp:90-117).  It converts a
// DOWorld into vg_scene_desc (include/vedo_gpu.h), steps on the GPU, and
// hands state and histories back to VEDO's own containers at the sync
// points the CPU path uses (ISReset / ISRecord, NBSConsultant.cpp:40-59),
// so WriteIDO and post-processing work unchanged.  Compiled against the
// reference's headers with its build macros (-D_UNIX_LIKE -D_INT32
// -D_FLOAT64 -D_STD_CPP_11) and linked with libvedo_hip.so.
#ifndef VEDO_AMD_GPU_SIM_MEDIATOR_H
#define VEDO_AMD_GPU_SIM_MEDIATOR_H

#include <aries/utility/Constants.h>
#include <vedo/framework/DOWorld.h>
#include <vedo/framework/IactRecordTab.h>

#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "vedo_gpu.h"

namespace vedo {

class GpuSimMediator {
public:
    // consultant: 0 LeapConsultant's candidate list, 1 NBSConsultant's
    GpuSimMediator(DOWorld *w, IactRecordTab *irt, unsigned record, unsigned upiact, int consultant = 0)
        : world(w), irt(irt), record(record)
    {
        const SystemParameter *sp = w->GetSystemParameter();
        std::vector<std::string> groups;
        std::map<std::string, int> midx;
        for (const DOModel *m : w->GetDOModel()) {
            const int g = group_index(groups, m->GetDOGroup());
            vg_model vm = {};
            vm.shape = (int)m->GetShapeType(); vm.behavior = (int)m->GetBehavior(); vm.scope = (int)m->GetScope();
            vm.group = g;
            vm.density = m->GetDensity(); vm.density_factor = m->GetDensityFactor();
            const Vector3df f = m->GetExternalForce();
            vm.fext[0] = f.x(); vm.fext[1] = f.y(); vm.fext[2] = f.z();
            const DOShapeAttributes a = m->GetShapeAttributes();
            if (m->GetShapeType() == Sphere) vm.attr[0] = a.sphere.radius;
            if (m->GetShapeType() == QuasiCylinder) {
                vm.attr[0] = a.quasicylinder.radius; vm.attr[1] = a.quasicylinder.height;
            }
            if (m->GetShapeType() == QuasiPlate) {
                vm.attr[0] = a.quasiplate.width; vm.attr[1] = a.quasiplate.height; vm.attr[2] = a.quasiplate.length;
            }
            if (m->GetShapeType() == QuasiPlateWithCircularHole) {
                vm.attr[0] = a.quasiplatewithcircularhole.width; vm.attr[1] = a.quasiplatewithcircularhole.height;
                vm.attr[2] = a.quasiplatewithcircularhole.length;
                vm.attr[3] = a.quasiplatewithcircularhole.holeradius;
                vm.attr[4] = a.quasiplatewithcircularhole.holexoffset;
                vm.attr[5] = a.quasiplatewithcircularhole.holeyoffset;
            }
            midx[m->GetDOName()] = (int)models.size();
            model_names.push_back(m->GetDOName());
            models.push_back(vm);
        }
        // dense group x group table, first match (DOWorld::GetIactModel)
        iact.assign(groups.size() * groups.size(), vg_iact{});
        for (size_t i = 0; i < groups.size(); ++i)
            for (size_t j = 0; j < groups.size(); ++j)
                if (const IactModel *im = w->GetIactModel(groups[i], groups[j]))
                    fill_law(*im, iact[i * groups.size() + j]);
        const std::vector<DOStatus *> &st = w->GetDOStatus();
        for (const DOStatus *s : st) {
            model_of.push_back(midx.at(s->GetDOName()));
            put(pos, s->GetPosition()); put(vel, s->GetVelocity()); put(ox, s->GetOrientationX());
            put(oz, s->GetOrientationZ()); put(av, s->GetAngularVelocity());
        }
        vg_scene_desc d = {};
        d.n_do = (int)st.size(); d.n_models = (int)models.size(); d.n_groups = (int)groups.size();
        d.models = models.data(); d.iact = iact.data(); d.model_of = model_of.data();
        d.pos = pos.data(); d.vel = vel.data(); d.ox = ox.data(); d.oz = oz.data(); d.av = av.data();
        d.dt = sp->GetTimeInterval(); d.t_start = sp->GetTimeStart(); d.t_stop = sp->GetTimeStop();
        d.t_current = sp->GetTimeCurrent();
        const Vector3df g = sp->GetFieldAcceleration();
        d.g[0] = g.x(); d.g[1] = g.y(); d.g[2] = g.z();
        d.safety_factor = aries::Constants::Instance()->SafetyFactor();
        d.up_iact = upiact; d.record = record; d.consultant = consultant;
        // Boundary ZoneOfInterest / PeriodicBoundaryConditions: an axis is
        // bounded / periodic where lower != upper
        bounds(sp->GetZoneOfInterest(), d.zoi_lo, d.zoi_hi);
        bounds(sp->GetPeriodicBoundaryConditions(), d.pbc_lo, d.pbc_hi);
        if (vg_create(&d, &h) != VG_OK) throw std::runtime_error(vg_create_error());
        // a restart: the IactRecordTab read with the .ido seeds the statuses (SimMediator::Initiate)
        if (irt->GetTabSize() > 0) {
            std::vector<unsigned> lo, hi;
            std::vector<vg_status> rs;
            for (const auto &e : irt->GetData()) {   // key (lo, hi) -> ImpactStatus
                lo.push_back((unsigned)e.first.first);
                hi.push_back((unsigned)e.first.second);
                const ImpactStatus &s = e.second;
                vg_status r = {};
                r.contact = s.Contact(); r.last_contact = s.LastContact(); r.bond = s.Bond();
                r.last_bond = s.LastBond(); r.kn = s.Kn(); r.init_vel = s.InitialVelocity();
                const Vector3df f = s.ShearForce();
                r.shear[0] = f.x(); r.shear[1] = f.y(); r.shear[2] = f.z();
                rs.push_back(r);
            }
            check(vg_set_history(h, (int)lo.size(), lo.data(), hi.data(), rs.data()));
        }
    }
    ~GpuSimMediator() { vg_destroy(h); }
    GpuSimMediator(const GpuSimMediator &) = delete;
    GpuSimMediator &operator=(const GpuSimMediator &) = delete;

    // Equivalent of `steps` calls of SimMediator::Run(); then the DOWorld
    // and the IactRecordTab hold the state a record point writes.
    bool Run(unsigned steps)
    {
        check(vg_step(h, (int)steps));
        check(vg_sync(h));
        SyncWorld();
        return vg_time(h) < world->GetSystemParameter()->GetTimeStop();
    }

    // DOWorld::UpdateDOStatus and RebuildIactRecordTab from the device
    void SyncWorld()
    {
        // the zone of interest may have erased DOs: the remaining ones, closed up, with their DONames
        int n = 0;
        check(vg_scene_num_do(h, 0, &n));
        const int n_old = (int)world->GetDOStatus().size();
        std::vector<int> mo(n);
        check(vg_scene_models(h, 0, mo.data()));
        if (n != n_old) {
            std::vector<_uint_t> tail;
            for (int i = n; i < n_old; ++i) tail.push_back((_uint_t)i);
            world->EraseDOStatus(tail);
        }
        std::vector<double> p(3 * n), v(3 * n), x(3 * n), z(3 * n), w(3 * n), I(3 * n), L(3 * n);
        check(vg_get_state(h, p.data(), v.data(), x.data(), z.data(), w.data(), I.data(), L.data()));
        for (int i = 0; i < n; ++i)
            world->SetDOStatus((_uint_t)i, DOStatus(model_names[mo[i]], V(p, i), V(v, i), V(x, i), V(z, i), V(w, i),
                                                    V(I, i), V(L, i)));
        long long m = 0;
        check(vg_get_active(h, &m, nullptr, nullptr, nullptr));
        std::vector<unsigned> lo(m), hi(m);
        std::vector<vg_status> rs(m);
        check(vg_get_active(h, &m, lo.data(), hi.data(), rs.data()));
        irt->Clear();
        for (long long k = 0; k < m; ++k) {
            const vg_status &r = rs[k];
            const bool c = r.contact != 0, b = r.bond != 0;
            const ImpactStatus s(c, b, r.kn, r.init_vel, Vector3df(r.shear[0], r.shear[1], r.shear[2]),
                                 Vector3df(r.point[0], r.point[1], r.point[2]),
                                 Vector3df(r.dir[0], r.dir[1], r.dir[2]),
                                 Vector3df(r.to_master[0], r.to_master[1], r.to_master[2]),
                                 Vector3df(r.ang_to_master[0], r.ang_to_master[1], r.ang_to_master[2]), r.overlap);
            irt->PushRecord((_uint_t)lo[k], (_uint_t)hi[k], s);   // std::map::insert: a key listed twice keeps its first
        }
    }

private:
    static int group_index(std::vector<std::string> &g, const std::string &s)
    {
        for (size_t k = 0; k < g.size(); ++k)
            if (g[k] == s) return (int)k;
        g.push_back(s);
        return (int)g.size() - 1;
    }
    // mechanism slots in the reference constructors' reading order (vedo_gpu.h)
    static void fill_law(const IactModel &im, vg_iact &out)
    {
        static const char *const names[6][8] = {
            {0},
            {"NormalStiffness", "ShearStiffness", "NormalDampingRatio", "ShearDampingRatio", "NormalBondStrength",
             "ShearBondStrength", "FrictionCoefficient"},
            {"NormalStiffness", "ShearStiffness", "NormalDampingRatio", "ShearDampingRatio", "NormalBondStrength",
             "ShearBondStrength", "FrictionCoefficient"},
            {"NormalStiffness", "ShearStiffness", "NormalDampingRatio", "ShearDampingRatio", "NormalBondStrength",
             "ShearBondStrength", "FrictionCoefficient"},
            {"HertzNormalStiffness", "EquivalentMass", "ShearStiffnessCoefficient1", "ShearStiffnessCoefficient2",
             "DampingRatio", "KineticFrictionCoefficient"},
            {"NormalStiffness", "NormalDampingCoefficient"}};
        static const char *const kinds[] = {"", "ISwBSDBF", "ISwBSDB2F", "ISwBtSDBF", "ISwH", "ISwHertz"};
        const std::string eq = im.GetEquationType();
        for (int l = 1; l <= 5; ++l)
            if (eq == kinds[l]) {
                out.law = l;
                for (int q = 0; q < 8 && names[l][q]; ++q) out.p[q] = im.GetIactMechanism(names[l][q]);
            }
    }
    static void bounds(const Boundary &b, double *lo, double *hi)
    {
        const Vector3df l = b.GetLowerPoint(), u = b.GetUpperPoint();
        lo[0] = l.x(); lo[1] = l.y(); lo[2] = l.z();
        hi[0] = u.x(); hi[1] = u.y(); hi[2] = u.z();
    }
    static void put(std::vector<double> &a, const Vector3df &v)
    {
        a.push_back(v.x()); a.push_back(v.y()); a.push_back(v.z());
    }
    static Vector3df V(const std::vector<double> &a, int i) { return Vector3df(a[3 * i], a[3 * i + 1], a[3 * i + 2]); }
    void check(int rc)
    {
        if (rc != VG_OK) throw std::runtime_error(vg_last_error(h));
    }

    DOWorld *world;
    IactRecordTab *irt;
    unsigned record;
    vg_scene *h = nullptr;
    std::vector<std::string> model_names;   // DOName per model index of vg_scene_desc
    std::vector<vg_model> models;
    std::vector<vg_iact> iact;
    std::vector<int> model_of;
    std::vector<double> pos, vel, ox, oz, av;
};

}  // namespace vedo

#endif
