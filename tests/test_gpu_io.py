"""Worlds, robots and meshes loaded from files run on the GPU engine with the
oracle's answers: a world XML (Klampt .rob robot with .off link meshes,
Klampt object files, an .off terrain under a scale and offset) and a URDF
robot, checked against the FP64 oracle on the same loaded data and against
the same scene built in memory; file meshes loaded more than once share one
device BVH (kg_geom_clone) and answer as separately built ones do."""
import os

import numpy as np
import pytest

import oracle
from klampt_amd import engine, io, robotsim, synthetic as S

pytestmark = pytest.mark.gpu
BAND = 1e-6


def _feas(r):
    """Feasibility flags from a batched call's result tuple."""
    return np.asarray(r[0]).astype(bool)


def _agree(got, ow, orb, Q):
    want = ow.feasible_batch(orb, Q)[0].astype(bool)
    bad = np.nonzero(got != want)[0]
    if len(bad):
        sep = ow.distance_batch(orb, Q[bad], signed=True)
        assert (np.abs(sep).min(axis=1) < BAND).all(), bad[:8]
    return want


def _write_world(d, sc, terrain):
    """The scene as files: arm .rob + link .off meshes, one Klampt object file
    per box (row-major T, 17 digits: exact), terrain .off placed by the XML."""
    files = []
    for l, m in enumerate(sc.robot.link_meshes):
        fn = os.path.join(d, "link%d.off" % l)
        io.save_mesh(fn, m)
        files.append(fn)
    rf = io.RobotFile(sc.robot, link_names=["l%d" % i for i in range(sc.robot.n)], q=np.zeros(sc.robot.n),
                      geom_files=files)
    io.save_rob(os.path.join(d, "arm.rob"), rf)
    xml = ['<world>', '  <robot name="arm" file="arm.rob"/>']
    for k, (mesh, xf) in enumerate(sc.objects):
        io.save_mesh(os.path.join(d, "box%d.off" % k), mesh)
        R, t = io.xf_mat(xf)
        with open(os.path.join(d, "box%d.obj" % k), "w") as f:
            f.write('mesh "box%d.off"\n' % k)
            f.write("T " + " ".join("%.17g" % v for v in list(R.reshape(9)) + list(t)) + "\n")
        xml.append('  <rigidObject name="box%d" file="box%d.obj"/>' % (k, k))
    io.save_mesh(os.path.join(d, "soup.off"), terrain)
    xml.append('  <terrain file="soup.off" scale="0.5" position="0.6 0 0.1"/>')
    xml.append('</world>')
    with open(os.path.join(d, "world.xml"), "w") as f:
        f.write("\n".join(xml) + "\n")
    return os.path.join(d, "world.xml")


@pytest.fixture(scope="module")
def loaded(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("world"))
    sc = S.boxes_scene(0)
    soup = S.clutter_soup(seed=5, target_tris=20_000, center=(0.8, 0.0, 0.4), side=1.0)
    path = _write_world(d, sc, soup)
    W = robotsim.WorldModel()
    assert W.loadFile(path)
    wf = io.load_world(path)
    ow, orb = oracle.world_from_scene(wf.scene())
    return sc, soup, W, wf, ow, orb


def test_loaded_world_feasibility_matches_oracle(loaded):
    sc, soup, W, wf, ow, orb = loaded
    assert W.numRobots() == 1 and W.numRigidObjects() == 16 and W.numTerrains() == 1
    space = robotsim.RobotCSpace(W.robot(0), W)
    Q = S.uniform_configs(sc.robot, 4096, seed=71)
    got = _feas(space.feasible_batch(Q))
    want = _agree(got, ow, orb, Q)
    assert 0 < want.sum() < len(Q)


def test_loaded_world_equals_in_memory_scene(loaded):
    """Same geometry bit for bit (17-digit text, exact rigid transforms), so
    the engine answers identically on both worlds."""
    sc, soup, W, wf, ow, orb = loaded
    mem = S.Scene(sc.robot, [io.transform_mesh(soup, np.array([[0.5, 0, 0, 0.6], [0, 0.5, 0, 0], [0, 0, 0.5, 0.1]]))],
                  list(sc.objects))
    assert np.array_equal(mem.terrains[0].verts, wf.terrains[0].mesh.verts)
    for (m1, x1), o in zip(mem.objects, wf.objects):
        assert np.array_equal(m1.verts, o.mesh.verts) and np.array_equal(x1, o.xf)
    ctx = robotsim.default_context()
    w2, r2 = engine.world_from_scene(ctx, mem)
    Q = S.uniform_configs(sc.robot, 4096, seed=72)
    a = _feas(robotsim.RobotCSpace(W.robot(0), W).feasible_batch(Q))
    b = _feas(w2.feasible_batch(Q))
    assert np.array_equal(a, b)


def test_shared_geometry_answers_as_separate_builds(tmp_path):
    """Eight rigid objects from one file: one BVH build, seven clones; the
    world answers as one built from eight separate meshes."""
    box = S.box_mesh([0.12, 0.08, 0.2])
    io.save_mesh(str(tmp_path / "b.off"), box)
    rng = np.random.default_rng(3)
    xfs = [S.f32r(io.xf12(S.random_rotation(rng), rng.uniform(-0.8, 0.8, 3))) for _ in range(8)]
    robotsim.clear_geometry_cache()
    W = robotsim.WorldModel()
    for k, x in enumerate(xfs):
        R, t = io.xf_mat(x)
        with open(str(tmp_path / ("o%d.obj" % k)), "w") as f:
            f.write('mesh "b.off"\nT ' + " ".join("%.17g" % v for v in list(R.reshape(9)) + list(t)) + "\n")
        assert W.loadRigidObject(str(tmp_path / ("o%d.obj" % k))) is not None
    assert len(robotsim._geom_cache) == 1
    spec = S.arm7()
    rob = robotsim.RobotModel.from_spec(spec)
    W.addRobot(rob)
    sc = S.Scene(spec, [], [(box, x) for x in xfs])
    ow, orb = oracle.world_from_scene(sc)
    Q = S.uniform_configs(spec, 2048, seed=73)
    got = _feas(robotsim.RobotCSpace(rob, W).feasible_batch(Q))
    _agree(got, ow, orb, Q)
    w2, r2 = engine.world_from_scene(robotsim.default_context(), sc)
    assert np.array_equal(got, _feas(w2.feasible_batch(Q)))
    # a clone keeps its own transform and margin
    g1, g2 = W.rigidObject(0).geometry(), W.rigidObject(1).geometry()
    g2.setCollisionMargin(0.01)
    assert g1.getCollisionMargin() == 0.0 and g2.getCollisionMargin() == 0.01


def test_geometry_loadfile_collides_matches_oracle(tmp_path):
    src = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "objects", "cube.off")
    a, b = robotsim.Geometry3D(), robotsim.Geometry3D()
    assert a.loadFile(src) and b.loadFile(src)
    assert not robotsim.Geometry3D().loadFile(str(tmp_path / "missing.off"))
    ow = oracle.OracleWorld()
    cube = io.load_mesh(src)
    gi = ow.add_mesh(cube)
    rng = np.random.default_rng(4)
    I = [1, 0, 0, 0, 1, 0, 0, 0, 1]
    for k in range(64):
        R = S.random_rotation(rng)
        t = rng.uniform(-1.6, 1.6, 3)
        a.setCurrentTransform(I, [0, 0, 0])
        b.setCurrentTransform(list(S.rot_colmajor(R)), list(t))
        want = ow.mesh_collide(gi, io.xf12(np.eye(3), np.zeros(3)), gi, io.xf12(R, t))
        got = a.collides(b)
        if got != want:
            d = ow.mesh_distance(gi, io.xf12(np.eye(3), np.zeros(3)), gi, io.xf12(R, t))[0]
            assert abs(d) < BAND, (k, got, want, d)


def test_urdf_robot_feasibility_matches_oracle(tmp_path):
    io.save_mesh(str(tmp_path / "cap.off"), S.capsule_mesh(0.05, 0.3, axis=0))
    link = ('<link name="{n}"><collision><origin xyz="0.15 0 0"/><geometry><mesh filename="cap.off"/>'
            '</geometry></collision></link>')
    joint = ('<joint name="j{k}" type="revolute"><parent link="{p}"/><child link="{c}"/>'
             '<origin xyz="{x} 0 0" rpy="{r} 0 0"/><axis xyz="0 0 1"/><limit lower="-2.5" upper="2.5"/></joint>')
    names = ["a%d" % i for i in range(6)]
    body = ['<robot name="snake"><link name="world"/>'] + [link.format(n=n) for n in names]
    body += [joint.format(k=0, p="world", c=names[0], x=0, r=0)]
    body += [joint.format(k=i, p=names[i - 1], c=names[i], x=0.3, r=0.5 * i) for i in range(1, 6)]
    body += ["</robot>"]
    p = tmp_path / "snake.urdf"
    p.write_text("\n".join(body))
    W = robotsim.WorldModel()
    rob = W.loadRobot(str(p))
    assert rob is not None and rob.numLinks() == 6 and rob.link("a3").getName() == "a3"
    box = S.box_mesh([0.3, 0.3, 0.05])
    W.addRigidObject(robotsim.Geometry3D(robotsim.TriangleMesh(box.verts, box.tris)), [1, 0, 0, 0, 1, 0, 0, 0, 1],
                     [0.6, 0, -0.2])
    rf = io.load_urdf(str(p))
    sc = S.Scene(rf.spec, [], [(box, io.xf12(np.eye(3), [0.6, 0, -0.2]))])
    ow, orb = oracle.world_from_scene(sc)
    Q = S.uniform_configs(rf.spec, 4096, seed=74)
    got = _feas(robotsim.RobotCSpace(rob, W).feasible_batch(Q))
    want = _agree(got, ow, orb, Q)
    assert 0 < want.sum() < len(Q)


def test_rigid_object_set_transform_moves_it(tmp_path):
    box = S.box_mesh([0.2, 0.2, 0.2])
    io.save_mesh(str(tmp_path / "b.off"), box)
    W = robotsim.WorldModel()
    obj = W.loadRigidObject(str(tmp_path / "b.off"))          # a mesh file: identity transform
    assert obj is not None and obj.getName() == "b" and obj.getID() == 0
    rob = robotsim.RobotModel.from_spec(S.arm7())
    W.addRobot(rob)
    space = robotsim.RobotCSpace(rob, W)
    Q = S.uniform_configs(S.arm7(), 512, seed=75)
    ow, orb = oracle.world_from_scene(S.Scene(S.arm7(), [], [(box, io.xf12(np.eye(3), np.zeros(3)))]))
    feas0 = _agree(_feas(space.feasible_batch(Q)), ow, orb, Q)
    assert not feas0.all()                                     # the box at the base blocks some configurations
    obj.setTransform([1, 0, 0, 0, 1, 0, 0, 0, 1], [10.0, 10.0, 10.0])
    assert obj.getTransform()[1] == [10.0, 10.0, 10.0]
    W2 = robotsim.WorldModel()
    rob2 = robotsim.RobotModel.from_spec(S.arm7())
    W2.addRobot(rob2)
    free = _feas(robotsim.RobotCSpace(rob2, W2).feasible_batch(Q))
    assert np.array_equal(_feas(space.feasible_batch(Q)), free) and free.any()


def test_loadfile_into_attached_geometry_updates_the_engine(tmp_path):
    """Loading a file into a link's or rigid object's Geometry3D replaces the
    geometry the world checks (robotsim.cpp:1443-1453), with the object's
    placement kept."""
    from klampt_amd import io as kio, robotsim
    spec = S.arm7()
    spec.link_meshes = [None] * 7
    W = robotsim.WorldModel()
    R = robotsim.RobotModel.from_spec(spec)
    W.addRobot(R)
    q = [0.0] * 7
    R.setConfig(q)
    Rl, tl = R.link(2).getTransform()
    # an object box centred on link 2's frame origin (placed by its transform)
    big = S.box_mesh([0.08, 0.08, 0.08])
    W.addRigidObject(robotsim.Geometry3D(robotsim.TriangleMesh(big.verts, big.tris)),
                     [1, 0, 0, 0, 1, 0, 0, 0, 1], list(np.asarray(tl) + [0.05, 0.0, 0.0]))
    space = robotsim.RobotCSpace(R, W)
    assert space.isFeasible(q)                       # the links have no geometry yet
    link_mesh = S.box_mesh([0.06, 0.06, 0.06])
    kio.save_mesh(str(tmp_path / "link.off"), link_mesh)
    assert R.link(2).geometry().loadFile(str(tmp_path / "link.off"))
    assert not space.isFeasible(q)                   # the loaded link mesh crosses the object
    # the object's geometry replaced by a small box: placement kept, no contact
    kio.save_mesh(str(tmp_path / "tiny.off"), S.box_mesh([0.005, 0.005, 0.005], c=[0.3, 0.3, 0.3]))
    assert W.rigidObject(0).geometry().loadFile(str(tmp_path / "tiny.off"))
    assert space.isFeasible(q)
    Ro, to = W.rigidObject(0).getTransform()
    assert np.allclose(to, np.asarray(tl) + [0.05, 0.0, 0.0])
