function [R,K,ok]=uvhAssemblyGPU(CtrlVar,MUA,F0,F1)
%uvhAssemblyGPU  R and K of the uvh system on the GPU through uvhAssemblyGPU_mex (libuvh).
%
%   [R,K,ok]=uvhAssemblyGPU(CtrlVar,MUA,F0,F1)
%
% Packs the arguments of uvhMatrixAssembly (UaMain/uvhMatrixAssembly.m:6) into the MEX
% gateway's plain arrays. ok is false when the options are outside the GPU library; the
% caller then assembles on the host with uvhMatrixAssembly.

persistent meshId lastMUA
ok=false; R=[]; K=[];

laws=["Weertman","W","Budd","W-N0","Coulomb","C","Tsai","minCW-N0","Cornford","rpCW-N0","Umbi","rCW-N0", ...
    "Joughin","rCW-v0"];
if CtrlVar.FlowApproximation~="SSTREAM" || ~ismember(CtrlVar.SlidingLaw,laws) ...
        || CtrlVar.IncludeMelangeModelPhysics || CtrlVar.uvhGroupAssembly ...
        || (CtrlVar.IncludeTG3uvhBoundaryTerm && CtrlVar.TG3) || CtrlVar.Development.Pre2025uvhAssembly ...
        || CtrlVar.OnlyCalcBasalDragAndEffectiveViscosity ...
        || (CtrlVar.MassBalanceGeometryFeedback>=2 && ~CtrlVar.uvhMatrixAssembly.ZeroFields) ...
        || (CtrlVar.LevelSetMethod && CtrlVar.LevelSetMethodAutomaticallyApplyMassBalanceFeedback) ...
        || (isfield(CtrlVar,'ThicknessPenalty') && CtrlVar.ThicknessPenalty) ...
        || (isfield(CtrlVar,'ThicknessBarrier') && CtrlVar.ThicknessBarrier)
    return
end

% a new id whenever the mesh changes (the gateway keeps the mesh and pattern on the device)
if isempty(meshId) || ~isequal(lastMUA.connectivity,MUA.connectivity) || ~isequal(lastMUA.coordinates,MUA.coordinates)
    meshId=now; lastMUA.connectivity=MUA.connectivity; lastMUA.coordinates=MUA.coordinates;
end

taus=["taus","taut","tau1","tau2","tau0"];
tau=find(taus==CtrlVar.uvh.SUPG.tau)-1;
ctrl=[CtrlVar.dt CtrlVar.theta CtrlVar.SUPG.beta0 CtrlVar.uvh.SUPG.tauMultiplier CtrlVar.SpeedZero ...
    CtrlVar.EpsZero CtrlVar.etaZero CtrlVar.Czero CtrlVar.kH CtrlVar.Hh0 F1.g F1.rhow F1.alpha ...
    CtrlVar.ThickMin tau (CtrlVar.ResetThicknessToMinThickness && CtrlVar.ResetThicknessInNonLinLoop) ...
    CtrlVar.Enforce_bAboveB floor((find(laws==CtrlVar.SlidingLaw,1)-1)/2) CtrlVar.Nzero];
Nn=MUA.Nnodes;
q=zeros(Nn,1); muk=zeros(Nn,1); V0=zeros(Nn,1);   % read only by the laws that use them
if ~isempty(F1.q), q=F1.q; end
if ~isempty(F1.muk), muk=F1.muk; end
if ~isempty(F1.V0), V0=F1.V0; end
fields=[F1.h F1.ub F1.vb F1.S F1.B F1.rho F1.C F1.m F1.n F1.AGlen F1.as F1.ab ...
    F0.h F0.ub F0.vb F0.as F0.ab zeros(Nn,1) q muk V0];
flags=double(CtrlVar.uvhMatrixAssembly.Ronly)+2*double(CtrlVar.uvhMatrixAssembly.ZeroFields);
if isfield(CtrlVar,'uvhGPU') && isfield(CtrlVar.uvhGPU,'FP32') && CtrlVar.uvhGPU.FP32
    flags=flags+4;
end

[R,K,status]=uvhAssemblyGPU_mex(meshId,MUA.connectivity,MUA.coordinates,MUA.points,MUA.weights,ctrl,fields,flags);
ok=(status==0);

end
